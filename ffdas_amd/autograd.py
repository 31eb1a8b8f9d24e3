"""Differentiable delay-and-sum: ``torch.autograd.Function`` over the hand-written adjoint.

Forward: ffdas_das_ex (tau = s * (offsets + r)).  Backward: ffdas_das_backward, which
recomputes every interpolation index from the geometry (nothing is stored between forward
and backward besides the inputs) and returns the gradients with respect to the channel data,
the transmit delays (offsets), the apodization weights and the delay scale s.  Gradients are
exact for the piecewise-linear interpolant (the right derivative at integer delays);
positions, directivity and sparse indices are not differentiable.
"""
from __future__ import annotations

import torch

from . import das as _das


class DasFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, offsets, weights, scale, srcpos, dstpos, srcdir, sparse_indices, wavenum,
                compute_type, channels_trailing, algorithm):
        s = float(scale.detach().item()) if isinstance(scale, torch.Tensor) else float(scale)
        x_, srcpos_, dstpos_, offsets_, weights_, srcdir_, si_, spatial, out_dims = _das._prepare(
            x, srcpos, dstpos, offsets, weights, srcdir, sparse_indices)
        out = _das._forward_call(x_, srcpos_, dstpos_, offsets_, weights_, srcdir_, wavenum, si_,
                                 s, algorithm, compute_type, channels_trailing, None, out_dims)
        ctx.save_for_backward(x, offsets, weights, srcpos, dstpos, srcdir, sparse_indices)
        ctx.s = s
        ctx.scale_is_tensor = isinstance(scale, torch.Tensor)
        ctx.wavenum = wavenum
        ctx.compute_type = compute_type
        ctx.channels_trailing = channels_trailing
        ctx.algorithm = algorithm
        return out

    @staticmethod
    def backward(ctx, grad_out):
        x, offsets, weights, srcpos, dstpos, srcdir, sparse_indices = ctx.saved_tensors
        need = ctx.needs_input_grad
        ct = ctx.compute_type
        gx, go, gw, gs = _das.das_backward(
            x, srcpos, dstpos, offsets, weights, grad_out, srcdir=srcdir, wavenum=ctx.wavenum,
            sparse_indices=sparse_indices, scale=ctx.s, use_fp16=(ct == _das.ComputeType.FP16),
            use_bf16=(ct == _das.ComputeType.BF16), channels_trailing=ctx.channels_trailing,
            need_x=need[0], need_offsets=need[1], need_weights=need[2],
            need_scale=need[3] and ctx.scale_is_tensor, algorithm=ctx.algorithm,
            reuse_x=True)   # x is a saved tensor: autograd rejects in-place changes to it
        if gx is not None and gx.dtype != x.dtype:
            gx = gx.to(x.dtype)
        if go is not None:
            go = go.reshape(offsets.shape).to(offsets.dtype)
        if gw is not None:
            gw = gw.reshape(weights.shape).to(weights.dtype)
        if gs is not None:
            gs = gs.to(torch.float32)
        return gx, go, gw, gs, None, None, None, None, None, None, None, None


def das_autograd(x, srcpos, dstpos, offsets, weights, *, scale=1.0, srcdir=None, wavenum=0.0,
                 sparse_indices=None, use_fp16=False, use_bf16=False, channels_trailing=False,
                 algorithm=_das.Algorithm.DEFAULT):
    """Differentiable das / das_sparse (always records the autograd graph)."""
    ct = _das._compute_type(_das.as_tensor(x), use_fp16, use_bf16)
    return DasFunction.apply(x, offsets, weights, scale, srcpos, dstpos, srcdir, sparse_indices,
                             float(wavenum), ct, bool(channels_trailing), algorithm)
