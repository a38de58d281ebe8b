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


class _Intensities(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V, conditions, kinetic, options, adjoint_options):
        table = BT.Table(V.detach().cpu().numpy(), kinetic)
        batch = BT.Batch(table, conditions)
        batch.solve(options)
        P = torch.from_numpy(batch.intensities_flat()).to(V.device)
        # the handles live as long as the graph node (ctx): a backward with retain_graph=True
        # can run again; they are released when ctx is freed
        ctx.batch, ctx.table = batch, table
        ctx.adjoint_options = adjoint_options or BT.SolveOptions(method="jacobi", precision="fp64",
                                                                 eps=FP64_ADJOINT_EPS)
        return P

    @staticmethod
    def backward(ctx, grad_P):
        g = grad_P.detach().to(torch.float64).cpu().numpy()
        grad = ctx.batch.adjoint(g, ctx.adjoint_options)
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
