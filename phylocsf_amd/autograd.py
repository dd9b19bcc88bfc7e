"""torch.autograd binding of the scoring path.

lpr(rho, batch) returns the per-region log-likelihoods lpr_leaves
(src/PhyloCSFModel.ml:67-82) of every model at per-region tree scales rho
[R, n_models]; its backward is pcsf_backward of the HIP library (the FP64
outside pass, or the FP32 tangent pass after an MFMA-precision forward), so
gradient-based fitting of the tree scale can use torch optimisers.  Forward keeps only the batch handle: the adjoint kernel
recomputes the inside pass.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import PCSF_FP32, PCSF_FP64


class _Lpr(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, batch, precision):
        r = rho.detach().to("cpu", torch.float64).numpy()
        ll, _ = batch.forward(r, precision, per_region=True)
        ctx.batch = batch
        ctx.rho = r
        ctx.precision = precision
        ctx.token = batch.forward_count
        return torch.from_numpy(ll).to(rho.device, rho.dtype)

    @staticmethod
    def backward(ctx, grad_ll):
        g = grad_ll.detach().to("cpu", torch.float64).numpy()
        if ctx.batch.forward_count != ctx.token:
            # the batch ran another forward since: differentiate at this
            # call's tree scales and precision, not the batch's latest
            ctx.batch.forward(ctx.rho, ctx.precision, per_region=True)
        grho, _ = ctx.batch.backward(g)
        return torch.from_numpy(grho).to(grad_ll.device, grad_ll.dtype), None, None


def lpr(rho: torch.Tensor, batch, precision=PCSF_FP64) -> torch.Tensor:
    """Per-region lpr [R, n_models] at tree scales rho [R, n_models], differentiable in rho."""
    return _Lpr.apply(rho, batch, precision)


__all__ = ["lpr", "PCSF_FP32", "PCSF_FP64", "np"]
