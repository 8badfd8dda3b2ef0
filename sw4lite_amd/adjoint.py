"""Receiver traces as a differentiable function of the material.

:class:`Seismograms` is a ``torch.autograd.Function`` over the C ABI: the
forward pass runs the fused FP64 step of an :class:`sw4lite_amd.EW` from zero
initial data and returns its receiver samples; the backward pass calls
``sw4hip_ew_gradient``, the discrete adjoint of the same step (transposed
kernels, checkpointed forward recomputation), and returns the gradients with
respect to rho, mu and lambda on every point of the grid.  The EW object must
be set up (sources, receivers, dt) before the first call; a new material is
loaded with ``set_material`` so dt stays the one of the set-up.
"""
import torch


class Seismograms(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, mu, la, ew, nsteps, ckpt):
        ew.set_material(rho.detach().contiguous(), mu.detach().contiguous(), la.detach().contiguous())
        ew.sync()
        ew.reset()
        ew.timesteploop(nsteps, mode=0)
        tr = torch.from_numpy(ew.receivers()[: nsteps + 1].copy()).to(rho.device)
        ctx.ew, ctx.nsteps, ctx.ckpt = ew, nsteps, ckpt
        ctx.save_for_backward(rho, mu, la)
        return tr

    @staticmethod
    def backward(ctx, grad):
        rho, mu, la = ctx.saved_tensors
        ew = ctx.ew
        ew.set_material(rho.detach().contiguous(), mu.detach().contiguous(), la.detach().contiguous())
        ew.sync()
        out = [torch.empty_like(rho), torch.empty_like(mu), torch.empty_like(la)]
        ew.gradient(grad.detach().to("cpu", torch.float64).numpy(), ctx.nsteps, ctx.ckpt, out=out)
        return out[0], out[1], out[2], None, None, None


def seismograms(rho, mu, la, ew, nsteps, ckpt=0):
    """Receiver samples [nsteps+1, nreceivers, 3] of the EW object for the
    material (rho, mu, la), differentiable in the three material fields."""
    return Seismograms.apply(rho, mu, la, ew, nsteps, ckpt)
