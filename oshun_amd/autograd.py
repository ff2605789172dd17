"""torch.autograd wrapper of the checkpointed adjoint (oshun_adjoint_forward /
oshun_adjoint_backward).

``run_steps(ctx, f, fld, h, nsteps, every)`` advances a 1D FP64 context with
m0 = 0 by ``nsteps`` steps of ``oshun_step`` starting from the device tensors
``f`` (ncomp x np x nx, the device-state layout) and ``fld`` (6 x nx), and
returns the final state as new device tensors.  Its backward pass is the
hand-written adjoint of the library: the states every ``every`` steps are kept
on the device and each segment is recomputed from its checkpoint.

A context records one run at a time: a second ``run_steps`` on the same
context replaces the first one's checkpoints, so the backward pass of the
first then raises OshunError instead of returning the gradient of the wrong
trajectory.  Chain segments through separate contexts, or call backward
before the next ``run_steps``.
"""
import math

import torch


class _Steps(torch.autograd.Function):
    @staticmethod
    def forward(actx, f, fld, ctx, h, nsteps, every):
        f = f.detach().contiguous()
        fld = fld.detach().contiguous()
        torch.cuda.synchronize(f.device)  # the library works on its own stream
        ctx.copy_device_state(f.data_ptr(), fld.data_ptr(), True)
        actx.generation = ctx.adjoint_forward(h, nsteps, every)
        f_out = torch.empty_like(f)
        fld_out = torch.empty_like(fld)
        ctx.copy_device_state(f_out.data_ptr(), fld_out.data_ptr(), False)
        actx.ctx = ctx
        actx.f_shape, actx.fld_shape = f.shape, fld.shape
        return f_out, fld_out

    @staticmethod
    def backward(actx, g_f, g_fld):
        shp, fshp = actx.ctx.device_shape(), actx.ctx.fld_device_shape()
        dev = (g_f if g_f is not None else g_fld).device
        lam = (g_f.detach().clone().contiguous() if g_f is not None
               else torch.zeros(shp, dtype=torch.float64, device=dev)).reshape(shp)
        lamf = (g_fld.detach().clone().contiguous() if g_fld is not None
                else torch.zeros(fshp, dtype=torch.float64, device=dev)).reshape(fshp)
        torch.cuda.synchronize(lam.device)
        actx.ctx.adjoint_backward(lam.data_ptr(), lamf.data_ptr(), actx.generation)
        return lam.reshape(actx.f_shape), lamf.reshape(actx.fld_shape), None, None, None, None


def run_steps(ctx, f, fld, h, nsteps, every):
    """Differentiable nsteps x oshun_step of a 1D m0 = 0 FP64 context."""
    for name, t, shape in (("f", f, ctx.device_shape()), ("fld", fld, ctx.fld_device_shape())):
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda:
            raise ValueError(f"run_steps: {name} must be a float64 tensor on the GPU")
        if t.numel() != math.prod(shape):
            raise ValueError(f"run_steps: {name} has {t.numel()} elements, the context needs {math.prod(shape)} "
                             f"(device layout {tuple(shape)})")
    return _Steps.apply(f, fld, ctx, float(h), int(nsteps), int(every))
