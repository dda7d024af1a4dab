"""
Differentiable Barnes interpolation for PyTorch-ROCm tensors.

The forward is the FP32C batch (``fastbarnes.batched.BarnesBatch``); the backward is
the hand-written adjoint in libfastbarnes.so: the line filters are self-adjoint and
the transpose of the bilinear injection is a bilinear gather, so one backward costs
one elementwise step, the same line passes as the forward, and a per-point gather.
Gradients flow to the values and, optionally, to the point positions.  The reference
has no gradient; the adjoint is derived from its forward (interpolation.py:328-367).

    batch = BarnesBatch([dict(sigma=1.0, x0=(-26, 34), step=1/32, size=(2400, 1200), n=100000)])
    out = barnes_torch(pts, val, batch)        # pts (n, dim) float64 cuda, val (n,) float64 cuda
    loss = 0.5 * torch.nansum((out - target) ** 2)
    loss.backward()                            # pts.grad, val.grad

``out`` is the batch's flat float32 output (fields at batch.out_offsets, each of shape
batch.shapes[f]); nodes beyond max_dist are NaN and receive no gradient.
"""
import torch

from fastbarnes.batched import BarnesBatch

__all__ = ['BarnesFunction', 'barnes_torch', 'BarnesBatch']


def _stream():
    from fastbarnes import _lib
    _lib.lib()
    if _lib.RUNTIME != 'torch':
        raise RuntimeError('libfastbarnes.so is bound to the system HIP runtime, not to the one of PyTorch '
                           '(FASTBARNES_SYSTEM_HIP=1?); torch streams and tensors cannot be passed to it')
    return torch.cuda.current_stream().cuda_stream


def _check(t, dtype, shape, name):
    if not t.is_cuda:
        raise RuntimeError('%s must be a GPU tensor' % name)
    if t.dtype != dtype:
        raise RuntimeError('%s must be %s' % (name, dtype))
    if tuple(t.shape) != tuple(shape):
        raise RuntimeError('%s has shape %s, expected %s' % (name, tuple(t.shape), tuple(shape)))


class BarnesFunction(torch.autograd.Function):
    """out = Barnes(pts, val) for a fixed plan; differentiable in pts and val."""

    @staticmethod
    def forward(ctx, pts, val, batch):
        if batch.precision != 'fp32c':
            raise RuntimeError('barnes_torch needs a plan of precision fp32c')
        pts = pts.contiguous()
        val = val.contiguous()
        _check(pts, torch.float64, (batch.total_points, batch.dim), 'pts')
        _check(val, torch.float64, (batch.total_points,), 'val')
        out = torch.empty(batch.total_out, dtype=torch.float32, device=val.device)
        wsave = torch.empty(batch.total_out, dtype=torch.float32, device=val.device)
        batch.forward(pts.data_ptr(), val.data_ptr(), out.data_ptr(), wsave.data_ptr(), _stream())
        ctx.batch = batch
        ctx.save_for_backward(pts, val, out, wsave)
        return out

    @staticmethod
    def backward(ctx, gout):
        pts, val, out, wsave = ctx.saved_tensors
        batch = ctx.batch
        gout = gout.to(torch.float32).contiguous()
        gval = torch.empty_like(val)
        gpts = torch.empty_like(pts) if ctx.needs_input_grad[0] else None
        batch.backward(pts.data_ptr(), val.data_ptr(), out.data_ptr(), wsave.data_ptr(), gout.data_ptr(),
                       gval.data_ptr(), gpts.data_ptr() if gpts is not None else None, _stream())
        return gpts, (gval if ctx.needs_input_grad[1] else None), None


def barnes_torch(pts, val, batch):
    """Differentiable batched Barnes interpolation (see module docstring)."""
    return BarnesFunction.apply(pts, val, batch)
