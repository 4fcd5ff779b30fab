"""The fermionic log-weight of a chain as a torch.autograd.Function.

``fermion_log_weight(x, chain)`` returns Re ln_cw at the fields x (a (Lt, Nb)
float64 tensor) with the gradient of pqmc_gradient attached, so that
``(bosonic - fermion_log_weight(x, chain)).backward()`` yields the force of a
hybrid-Monte-Carlo move.  No counterpart in the reference.  The forward
leaves the chain at x after init_g(1) and a gradient pass, i.e. where a
backward sweep that accepts nothing leaves it; the GPU does all the work,
torch only carries the result.
"""
from __future__ import annotations

import numpy as np
import torch


class FermionLogWeight(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, chain) -> torch.Tensor:
        xs = np.ascontiguousarray(x.detach().to("cpu", torch.float64).numpy())
        chain.set_fields(xs)
        chain.init_g(1)
        w = chain.ln_cw
        g = chain.gradient().real
        ctx.save_for_backward(torch.from_numpy(g).to(x.device, x.dtype))
        return x.new_tensor(w.real)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (g,) = ctx.saved_tensors
        return grad_out * g, None


def fermion_log_weight(x: torch.Tensor, chain) -> torch.Tensor:
    return FermionLogWeight.apply(x, chain)
