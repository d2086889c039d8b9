"""PyTorch-ROCm autograd wrapper over the C ABI (forward + adjoint).

``icm_equity(S, payout, Q)`` returns the equities as a differentiable
function of the stacks S: the backward pass calls the library's adjoint
(``icm_hip_vjp_with_plan_device``) with the incoming gradient as the
cotangent, so d(v . E)/dS comes from one GPU pass over the same tree.  The
payout is treated as a constant (no gradient), and the quadrature nodes are
held fixed (they depend on max(S) only through the rule's upper limit).

Tensors stay on the GPU: the plan copies S once at creation (it needs the
stack order on the host), payout / cotangent / outputs are passed as device
pointers of float64 CUDA tensors.

Memory: the plan lives with the autograd graph (its adjoint buffers are
added at the first backward).  mem_fraction caps the share of free device
memory the plan plans for (default 0.6 of what is free at each allocation);
a plan that does not fit halves its q-batch, at most four times, before the
call fails.
"""
import torch

import icm_amd


class ICMEquity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, S, payout, Q=256, mem_fraction=0.0):
        if S.dtype != torch.float64 or payout.dtype != torch.float64:
            raise TypeError("icm_equity works in float64")
        if S.dim() != 1 or payout.dim() != 1:
            raise ValueError("S and payout must be 1-D")
        dev = S.device if S.is_cuda else torch.device("cuda:0")
        plan = icm_amd.Plan(S.detach().cpu().numpy(), payout.numel(), mem_fraction=mem_fraction)
        d_p = payout.detach().to(dev).contiguous()
        eq = torch.empty(S.numel(), dtype=torch.float64, device=dev)
        torch.cuda.synchronize(dev)
        plan.run_device(d_p.data_ptr(), eq.data_ptr(), int(Q))
        ctx.plan = plan
        ctx.Q = int(Q)
        ctx.save_for_backward(d_p)
        ctx.s_device = S.device
        return eq if S.is_cuda else eq.cpu()

    @staticmethod
    def backward(ctx, grad_eq):
        (d_p,) = ctx.saved_tensors
        dev = d_p.device
        v = grad_eq.detach().to(dev, torch.float64).contiguous()
        ds = torch.empty_like(v)
        torch.cuda.synchronize(dev)
        # the plan stays alive with ctx (repeated backward calls, e.g. a
        # Jacobian row by row, reuse it); it is freed with the graph
        ctx.plan.vjp_device(d_p.data_ptr(), v.data_ptr(), 0, ds.data_ptr(), ctx.Q)
        return ds.to(ctx.s_device), None, None, None


def icm_equity(S, payout, Q=256, mem_fraction=0.0):
    """Differentiable ICM equities (float64 tensors); gradient w.r.t. S only."""
    return ICMEquity.apply(S, payout, Q, mem_fraction)
