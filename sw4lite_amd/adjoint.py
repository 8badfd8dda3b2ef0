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
        if not (rho.is_cuda and mu.is_cuda and la.is_cuda):
            raise ValueError("rho, mu and lambda must be device tensors")
        f64 = [x.detach().to(torch.float64).contiguous() for x in (rho, mu, la)]
        ew.set_material(*f64)
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
        ew.set_material(*[x.detach().to(torch.float64).contiguous() for x in (rho, mu, la)])
        ew.sync()
        n = ew.npoints()
        out = [torch.empty(n, dtype=torch.float64, device=rho.device) for _ in range(3)]
        ew.gradient(grad.detach().to("cpu", torch.float64).numpy(), ctx.nsteps, ctx.ckpt, out=out)
        return (out[0].view_as(rho).to(rho.dtype), out[1].view_as(mu).to(mu.dtype), out[2].view_as(la).to(la.dtype),
                None, None, None)


def seismograms(rho, mu, la, ew, nsteps, ckpt=0):
    """Receiver samples [nsteps+1, nreceivers, 3] of the EW object for the
    material (rho, mu, la), differentiable in the three material fields."""
    return Seismograms.apply(rho, mu, la, ew, nsteps, ckpt)
