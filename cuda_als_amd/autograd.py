"""Differentiable ALS half-sweep for PyTorch (test harness and lambda tuning).

    X = HalfSweep.apply(lam, Y, engine, side)

solves, for every row u of `side`, x_u = A_u^-1 b_u with
A_u = sum y_i y_i^T + lam_u n_u I and b_u = sum r_ui y_i, on the GPU engine
(als_forward_dev_ws). Its backward is the engine's hand-written adjoint
(als_adjoint_dev_ws): xbar = A^-1 g on the saved Cholesky factor,
dL/dlam_u = -n_u xbar_u . x_u, dL/dY through A and b. Every forward call
keeps its own factorisation in a tensor of the autograd context, so
half-sweeps chained in one graph (users then items, unrolled sweeps) each
differentiate against their own factor.

Tensors: lam (rows,), Y (cols, f) float32 on the engine's device; X is
(rows, f). The engine works on a padded row stride (als_info.ld); padding is
done here. Run under a non-default torch stream: the engine is pointed at
torch's current stream so both order their work on it.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

from . import _capi as A


class HalfSweep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lam, Y, engine, side):
        ld = engine.info().ld
        f = engine.cfg.f
        rows = engine.m if side == A.ALS_SIDE_USERS else engine.n
        stream = torch.cuda.current_stream()
        if stream.cuda_stream == 0:
            raise RuntimeError("HalfSweep needs a non-default torch stream (torch.cuda.stream(...))")
        engine.set_stream(stream.cuda_stream)
        Yp = F.pad(Y.detach().float(), (0, ld - f)).contiguous()
        lamc = lam.detach().float().contiguous()
        Xp = torch.zeros(rows, ld, device=Y.device, dtype=torch.float32)
        nbytes = engine.factor_save_bytes(side)
        Ls = torch.empty(nbytes // 4, device=Y.device, dtype=torch.float32)
        engine.forward_dev_ws(side, Yp.data_ptr(), Xp.data_ptr(), lamc.data_ptr(), Ls.data_ptr(), nbytes)
        ctx.engine, ctx.side, ctx.f, ctx.ld = engine, side, f, ld
        ctx.save_for_backward(lamc, Yp, Xp, Ls)
        return Xp[:, :f].clone()

    @staticmethod
    def backward(ctx, gX):
        lamc, Yp, Xp, Ls = ctx.saved_tensors
        engine, side, f, ld = ctx.engine, ctx.side, ctx.f, ctx.ld
        stream = torch.cuda.current_stream()
        engine.set_stream(stream.cuda_stream)
        g = F.pad(gX.float(), (0, ld - f)).contiguous()
        xbar = torch.zeros_like(Xp)
        lbar = torch.zeros_like(lamc)
        ybar = torch.zeros_like(Yp)
        engine.adjoint_dev_ws(side, Yp.data_ptr(), Xp.data_ptr(), Ls.data_ptr(), Ls.numel() * 4, g.data_ptr(),
                              xbar.data_ptr(), lbar.data_ptr(), ybar.data_ptr())
        return lbar, ybar[:, :f], None, None


def half_sweep(lam: torch.Tensor, Y: torch.Tensor, engine, side: int = A.ALS_SIDE_USERS) -> torch.Tensor:
    return HalfSweep.apply(lam, Y, engine, side)
