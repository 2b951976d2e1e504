"""Differentiable PDA chi2 for PyTorch: the Rao-Blackwellised forward and the
hand-written adjoint kernel of libgsmfretda_pda.so as a torch.autograd.Function.

    M = RBExpectedHistogram.apply(args, mc, n, R, seed, set_id, sid)
    loss = chi2(M, observed, R, n)
    loss.backward()          # args.grad = d chi2 / d args via pda_rb_adjoint

`args` is the mc::set_params vector (src/mc.cu:790-800): E (n), the
off-diagonal rates K (n (n - 1), genMatK's row-major order), v (n, held at
0).  The expected histogram M is computed on the GPU; chi2 is the reference's
reduced chi2 (src/streamWorker.cpp:197-206) on M / R, written in torch so that
autograd supplies the cotangent d chi2 / d M the adjoint kernel contracts.
The walk's random streams are keyed by (seed, set_id), so evaluations at
different parameters share common random numbers.
"""
import numpy as np
import torch

from . import pda


class RBExpectedHistogram(torch.autograd.Function):
    """args -> expected E histogram M (float64, one entry per bin)."""

    @staticmethod
    def forward(ctx, args, mc, n, R, seed, set_id=0, sid=0):
        a = args.detach().to("cpu", torch.float64).numpy().astype(np.float32)
        if not mc.set_params(n, sid, a):
            raise pda.PdaError(f"bad parameter set {a}")
        mc.rb_forward(R=R, seed=seed, set_id=set_id, sid=sid, scores=True)
        M, _, _, _ = mc.rb_result(sid)
        ctx.mc, ctx.n, ctx.sid, ctx.token = mc, n, sid, mc.rb_run_id(sid)
        ctx.dtype, ctx.device = args.dtype, args.device
        return torch.from_numpy(M).to(args.device)

    @staticmethod
    def backward(ctx, gM):
        if ctx.mc.rb_run_id(ctx.sid) != ctx.token:
            raise pda.PdaError(f"another RB forward ran on slot {ctx.sid} since this one; "
                               "call backward before the next forward on that slot")
        g = gM.detach().to("cpu", torch.float64).numpy()
        grad = ctx.mc.rb_adjoint(g, ctx.n, ctx.sid)
        return (torch.from_numpy(grad).to(dtype=ctx.dtype, device=ctx.device),
                None, None, None, None, None, None)


def chi2(M, observed, R, n, left_pad=2):
    """Reduced chi2 of expected counts M against the observed histogram:
    sum over bins b >= left_pad with M_b > 0 of (O_b - M_b/R)^2 / (M_b/R),
    divided by (bins kept - n (n + 1))."""
    observed = torch.as_tensor(observed, dtype=M.dtype, device=M.device)
    m = M / R
    keep = (torch.arange(M.shape[0], device=M.device) >= left_pad) & (M > 0)
    safe = torch.where(keep, m, torch.ones_like(m))
    terms = torch.where(keep, (observed - m) ** 2 / safe, torch.zeros_like(m))
    dof = int(keep.sum()) - n * (n + 1)
    return terms.sum() / dof


def rb_chi2(args, mc, n, R, seed, observed, set_id=0, sid=0, left_pad=2):
    """chi2(args) with gradient: the RB forward, chi2 in torch, the adjoint on backward."""
    M = RBExpectedHistogram.apply(args, mc, n, R, seed, set_id, sid)
    return chi2(M, observed, R, n, left_pad)
