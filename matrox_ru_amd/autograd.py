"""torch.autograd over the C-ABI evaluator.

``hgemm_apply(plan, W, D, B, VT)`` returns Y = K~ W (leaf-blocked, flat
N*nrhs) and differentiates with respect to W and, when they are given as
tensors requiring grad, the stored blocks D, B, VT (FP64 device tensors in
the plan's CDS layout).  The backward pass is mx_backward: the adjoint
evaluation dW = K~^T dY and the NT block-gradient products.

The reference has no derivatives; the blocks and their layout are those of
matroxTest-V1/hcds.h:12-34 (D/B/VT, Dptr/Bptr/VTptr).
"""
import torch

_SKEL_DTYPE = {"fp64": torch.float64, "fp32": torch.float32, "bf16x3": torch.float32}


class HGEMMFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, W, D, B, VT, plan, nrhs, precision):
        blocks = (D, B, VT)
        if any(t is not None for t in blocks):
            plan.set_blocks(*[t.detach() if t is not None else None for t in blocks])
        W = W.detach().contiguous()
        Y = torch.empty_like(W)
        plan.eval(W, Y, nrhs, precision)
        need_blocks = tuple(bool(ctx.needs_input_grad[i]) for i in (1, 2, 3))
        ws = us = None
        if any(need_blocks):
            ws, us = plan.skeletons(nrhs, _SKEL_DTYPE[precision])
        ctx.plan, ctx.nrhs, ctx.precision, ctx.need_blocks = plan, nrhs, precision, need_blocks
        ctx.blocks = blocks
        ctx.save_for_backward(W, ws, us)
        return Y

    @staticmethod
    def backward(ctx, dY):
        W, ws, us = ctx.saved_tensors
        plan, nrhs, precision = ctx.plan, ctx.nrhs, ctx.precision
        dY = dY.contiguous().to(W.dtype)
        if any(t is not None for t in ctx.blocks):   # the plan may hold other blocks by now
            plan.set_blocks(*[t.detach() if t is not None else None for t in ctx.blocks])
        if any(ctx.need_blocks):
            dW, dD, dB, dVT = plan.backward(W, dY, nrhs, precision, skeletons=(ws, us), blocks=ctx.need_blocks)
            outs = []
            for g, t in zip((dD, dB, dVT), ctx.blocks):
                outs.append(g.to(t.dtype) if g is not None else None)
        else:
            dW = torch.empty_like(dY)
            plan.eval_adjoint(dY, dW, nrhs, precision)
            outs = [None, None, None]
        if not ctx.needs_input_grad[0]:
            dW = None
        return dW, outs[0], outs[1], outs[2], None, None, None


def hgemm_apply(plan, W, D=None, B=None, VT=None, nrhs=None, precision="fp64"):
    """Differentiable Y = K~ W on a Plan built with adjoint=True."""
    if nrhs is None:
        nrhs = W.numel() // plan.rows
    return HGEMMFunction.apply(W, D, B, VT, plan, nrhs, precision)
