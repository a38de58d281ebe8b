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


def test_autograd_reuses_its_batch_across_calls_and_reads_v_on_the_gpu():
    """Calls with the same conditions list share one table and batch (the potential replaced
    in place, from device memory when V lives on the GPU): a descent step's gradients equal
    those of fresh tables, and a backward whose forward state a later forward replaced solves
    its forward again."""
    w = W.hex_batch(n_conditions=3, nz=24, nk=7, cond_seed=11, channel_counts=[12, 20, 30])
    opts = BT.SolveOptions(method="jacobi", eps=5e-14)
    conds = list(w.conditions)

    def fresh_grad(Vn, target):
        tab = BT.Table(Vn, w.kinetic)
        bt = BT.Batch(tab, conds)
        bt.solve(opts)
        return bt.adjoint(2.0 * (bt.intensities_flat() - target), opts), bt.intensities_flat()

    rng = np.random.default_rng(4)
    V0 = w.potential.copy()
    _, P0 = fresh_grad(V0, np.zeros(1))
    target = P0 * (1 + 0.05 * rng.standard_normal(P0.size))
    for dev in ("cpu", "cuda"):
        Vn = V0.copy()
        slots = set()
        for step in range(3):
            V = torch.tensor(Vn, dtype=torch.complex128, device=dev, requires_grad=True)
            P = AG.scattering_intensities(V, conds, w.kinetic, opts, adjoint_options=opts)
            slots.add(id(AG._SLOT[0]))
            loss = ((P - torch.from_numpy(target).to(dev)) ** 2).sum()
            loss.backward()
            g = V.grad.cpu().numpy()
            ref, Pref = fresh_grad(Vn, target)
            np.testing.assert_allclose(P.detach().cpu().numpy(), Pref, rtol=0, atol=1e-13)
            np.testing.assert_allclose(g, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
            Vn = Vn - 1e-3 * g / np.abs(g).max() * np.abs(V0).max()
        assert len(slots) == 1, slots
    # two forwards, then the first one's backward: its state was replaced, so it solves again
    Va = torch.tensor(V0, dtype=torch.complex128, requires_grad=True)
    Vb = torch.tensor(V0 * 1.01, dtype=torch.complex128, requires_grad=True)
    la = ((AG.scattering_intensities(Va, conds, w.kinetic, opts, adjoint_options=opts)
           - torch.from_numpy(target)) ** 2).sum()
    lb = ((AG.scattering_intensities(Vb, conds, w.kinetic, opts, adjoint_options=opts)
           - torch.from_numpy(target)) ** 2).sum()
    la.backward()
    lb.backward()
    for Vt, Vn in ((Va, V0), (Vb, V0 * 1.01)):
        ref, _ = fresh_grad(Vn, target)
        np.testing.assert_allclose(Vt.grad.numpy(), ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
