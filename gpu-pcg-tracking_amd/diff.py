"""torch.autograd wrapper of the differentiable QP layer (C-ABI mpc_batch_qp_*).

TrackingLoss maps the per-dimension cost weights (wx: 14, wu: 7) and the
reference points ee_ref (6 floats per knot, xyz used) to the per-instance
tracking loss L_i = sum_k |ee(q*_k) - ref_k|^2 of the full-step solution
x* = xu - dz of the SQP step's QP at a fixed iterate xu (normally the final
iterate of mpc_batch_sqp).  The backward pass is the hand-written adjoint:
the linearisation is recomputed at xu and the same Schur complement,
block-Jacobi preconditioner and PCG solve K v = [-dL/dx*; 0].  Both solves
take `refine` steps of iterative refinement with an fp64 KKT residual.
"""
from __future__ import annotations

import numpy as np
import torch


class Layout:
    """Per-element instance index of a batch layout (for scaling by dL/dL_i)."""

    def __init__(self, horizons, device="cuda"):
        h = np.asarray(horizons, np.int64)
        self.n = len(h)
        self.knot_inst = torch.from_numpy(np.repeat(np.arange(self.n), h)).to(device)
        self.xu_inst = torch.from_numpy(np.repeat(np.arange(self.n), 21 * h - 7)).to(device)


class TrackingLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, wx, wu, ee_ref, solver, layout, xu, rho, lam, pcg_tol, refine):
        solver.set_cost_weights(wx.detach().cpu().numpy(), wu.detach().cpu().numpy())
        ee = ee_ref.detach().contiguous()
        dz, xstar = torch.empty_like(xu), torch.empty_like(xu)
        solver.qp_layer(xu, ee, rho, lam, dz, xstar, pcg_tol, refine)
        loss = torch.empty(layout.n, device=xu.device)
        dLdx, dLdref = torch.empty_like(xu), torch.empty_like(ee)
        solver.qp_loss(xstar, ee, loss, dLdx, dLdref)
        ctx.save_for_backward(wx, wu, ee, xu, rho, dz, dLdx, dLdref)
        ctx.solver, ctx.layout, ctx.pcg_tol, ctx.refine = solver, layout, pcg_tol, refine
        return loss

    @staticmethod
    def backward(ctx, gloss):
        wx, wu, ee, xu, rho, dz, dLdx, dLdref = ctx.saved_tensors
        solver, layout = ctx.solver, ctx.layout
        solver.set_cost_weights(wx.detach().cpu().numpy(), wu.detach().cpu().numpy())
        gloss = gloss.contiguous()
        a = (dLdx * gloss[layout.xu_inst]).contiguous()
        grad_w = torch.empty(layout.n * 21, device=xu.device)
        grad_ref = (dLdref.view(-1, 6) * gloss[layout.knot_inst].unsqueeze(1)).reshape(-1).contiguous()
        solver.qp_adjoint(xu, ee, rho, dz, a, grad_w, grad_ref, ctx.pcg_tol, ctx.refine)
        gw = grad_w.view(layout.n, 21).sum(0)
        return (gw[:14].to(wx.device, wx.dtype), gw[14:].to(wu.device, wu.dtype), grad_ref.view_as(ee),
                None, None, None, None, None, None, None)


def tracking_loss(solver, layout, wx, wu, ee_ref, xu, rho, lam, pcg_tol=0.0, refine=None):
    """Per-instance tracking loss of the QP layer at xu; differentiable in wx, wu and ee_ref."""
    if refine is None:
        from . import REFINE_STEPS as refine
    return TrackingLoss.apply(wx, wu, ee_ref, solver, layout, xu, rho, lam, pcg_tol, refine)
