"""Differentiable transforms for PyTorch over the library's device entry points.

``fftn(x, rank, weight=None, inverse=False)`` returns ``weight * F(x)`` where
F is the forward DFT over the last ``rank`` dims (or the unnormalised
backward DFT with ``inverse=True``) and ``weight`` is an optional real FP32
tensor of the transform's shape, broadcast over the batch.

Backward is the hand-written adjoint: for y = W . F x the gradient is
F^H (W . g), the unnormalised opposite-direction transform with W fused into
its first pass's loads (fftamd_c2c_weighted_dev), so the backward costs one
transform and no extra elementwise pass.  PyTorch's convention for complex
inputs is followed (grad = 2 dL/d conj(x) for a real loss).  A gradient for
``weight`` (Re sum_batch conj(g) . F x) is produced when it requires grad.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import last_error, load


def _stream(t):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _check(x, rank):
    if not x.is_cuda or x.dtype != torch.complex64:
        raise TypeError("complex64 CUDA/HIP tensor expected")
    if rank < 1 or rank > 3 or x.dim() < rank:
        raise ValueError("bad rank")


def weighted_transform(x, rank, inverse=False, w_in=None, w_out=None, out=None):
    """out = w_out . F(w_in . x) on the current stream (no autograd)."""
    _check(x, rank)
    x = x.contiguous()
    dims = [int(d) for d in x.shape[x.dim() - rank:]]
    batch = x.numel() // max(1, int(torch.tensor(dims).prod().item()))
    if out is None:
        out = torch.empty_like(x)
    for w in (w_in, w_out):
        if w is not None and (w.dtype != torch.float32 or not w.is_cuda or list(w.shape) != dims
                              or not w.is_contiguous()):
            raise ValueError("weights must be contiguous float32 CUDA/HIP tensors of the transform shape")
    if batch == 0:
        return out
    cd = (ctypes.c_uint * rank)(*dims)
    rc = load().fftamd_c2c_weighted_dev(
        rank, cd, batch, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
        1 if inverse else 0,
        ctypes.c_void_p(w_in.data_ptr()) if w_in is not None else None,
        ctypes.c_void_p(w_out.data_ptr()) if w_out is not None else None,
        _stream(x))
    if rc != 0:
        raise RuntimeError(f"fftamd weighted transform failed: {last_error()}")
    return out


class _WeightedFFT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, weight, rank, inverse):
        y = weighted_transform(x, rank, inverse=inverse, w_out=weight)
        ctx.rank = rank
        ctx.inverse = inverse
        ctx.has_w = weight is not None
        if weight is None:
            ctx.save_for_backward()
        elif weight.requires_grad:
            ctx.save_for_backward(weight, x)
        else:
            ctx.save_for_backward(weight)
        return y

    @staticmethod
    def backward(ctx, g):
        saved = ctx.saved_tensors
        w = saved[0] if ctx.has_w else None
        x = saved[1] if len(saved) > 1 else None
        g = g.contiguous()
        gx = None
        if ctx.needs_input_grad[0]:
            # adjoint of W . F is F^H . W: opposite direction, weights on the load
            gx = weighted_transform(g, ctx.rank, inverse=not ctx.inverse, w_in=w)
        gw = None
        if ctx.has_w and ctx.needs_input_grad[1]:
            fx = weighted_transform(x, ctx.rank, inverse=ctx.inverse)
            prod = (torch.conj(g) * fx).real
            gw = prod.reshape(-1, *w.shape).sum(0)
        return gx, gw, None, None


def fftn(x, rank: int = 3, weight=None, inverse: bool = False):
    """Differentiable weight . F(x) over the last `rank` dims (see module doc)."""
    return _WeightedFFT.apply(x, weight, rank, inverse)


def fft3(x, weight=None):
    return fftn(x, 3, weight, False)


def ifft3_unnormalised(x, weight=None):
    return fftn(x, 3, weight, True)
