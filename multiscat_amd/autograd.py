"""torch.autograd bridge: intensities of a batch of conditions as a differentiable function
of the potential table, with the backward pass done by the GPU adjoint solve.

    P = scattering_intensities(V, conditions, kinetic, options)   # V: complex128 tensor
    loss = ((P - P_target) ** 2).sum(); loss.backward()            # V.grad = dL/dRe V + i dL/dIm V

The gradient follows PyTorch's convention for a real loss of a complex tensor (the
conjugate Wirtinger derivative times two, i.e. dL/dRe V + i dL/dIm V), which is what
msx_batch_adjoint returns; V_00 enters only through Re V_00, so Im of its gradient is 0.
The forward solve and the adjoint run in libmultiscat_amd; torch only carries the arrays.
"""

from __future__ import annotations

import numpy as np
import torch

from multiscat_amd import batch as BT

# tolerance of the default adjoint solve (FP64 block Jacobi): far below the 1e-5 relative
# accuracy asked of gradients, so the adjoint's own error does not show in them
FP64_ADJOINT_EPS = 5e-13


class _Slot:
    """A table and a batch kept between calls for one conditions list (held here, so its id
    cannot be reused while the slot lives), one table shape and one kinetic matrix.
    `generation` counts the forward solves the batch has held."""

    def __init__(self, conditions, V, kinetic):
        self.conditions = conditions
        self.kinetic = kinetic
        self.table = BT.Table(np.zeros(tuple(V.shape), np.complex128), kinetic)
        self.batch = BT.Batch(self.table, conditions)
        self.generation = 0

    def matches(self, conditions, V, kinetic) -> bool:
        return (self.conditions is conditions and tuple(self.table.shape) == tuple(V.shape)
                and self.kinetic.shape == kinetic.shape and np.array_equal(self.kinetic, kinetic))

    def solve(self, V, options) -> int:
        self.table.set_potential(V)
        self.batch.solve(options)
        self.generation += 1
        return self.generation


# the most recent slot: an optimisation loop calls forward and backward with the same
# conditions and a new V, which then costs no table or batch set-up, and no upload of V
# through the host when V lives on the GPU. A backward whose forward state a later forward
# has replaced solves its forward again first.
_SLOT: list = []


def _slot_for(conditions, V, kinetic) -> _Slot:
    if _SLOT and _SLOT[0].matches(conditions, V, kinetic):
        return _SLOT[0]
    _SLOT[:] = [_Slot(conditions, V, kinetic)]
    return _SLOT[0]


class _Intensities(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V, conditions, kinetic, options, adjoint_options):
        slot = _slot_for(conditions, V, kinetic)
        ctx.generation = slot.solve(V.detach(), options)
        P = torch.from_numpy(slot.batch.intensities_flat()).to(V.device)
        # the slot lives as long as the graph node (ctx) needs it: a backward with
        # retain_graph=True can run again
        ctx.slot, ctx.V, ctx.options = slot, V.detach(), options
        ctx.adjoint_options = adjoint_options or BT.SolveOptions(method="jacobi", precision="fp64",
                                                                 eps=FP64_ADJOINT_EPS)
        return P

    @staticmethod
    def backward(ctx, grad_P):
        slot = ctx.slot
        if slot.generation != ctx.generation:
            ctx.generation = slot.solve(ctx.V, ctx.options)
        g = grad_P.detach().to(torch.float64).cpu().numpy()
        grad = slot.batch.adjoint(g, ctx.adjoint_options)
        return torch.from_numpy(grad).to(grad_P.device), None, None, None, None


def scattering_intensities(V: torch.Tensor, conditions, kinetic: np.ndarray,
                           options: BT.SolveOptions | None = None,
                           adjoint_options: BT.SolveOptions | None = None) -> torch.Tensor:
    """P for every channel of every condition (concatenated in batch order) as a function of
    the natural-unit table V (nkx, nky, nz) complex128. `conditions` are what
    multiscat_amd.batch.Batch takes; the adjoint solve uses `adjoint_options`, by default FP64
    block Jacobi to eps 5e-13 whatever the forward precision (a mixed-precision adjoint is
    available by passing mixed options)."""
    if V.dtype != torch.complex128:
        raise TypeError("V must be complex128")
    return _Intensities.apply(V, conditions, np.ascontiguousarray(kinetic, dtype=np.float64),
                              options or BT.SolveOptions(), adjoint_options)
