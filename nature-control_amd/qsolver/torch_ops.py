"""torch.autograd.Function wrappers of the fused IFRK2 step and its discrete
adjoint (qs_adj_* over the C-ABI).  No reference counterpart: the reference
has no adjoint.

Complex gradients follow torch's convention for a real loss: the gradient of a
complex tensor z is dL/dRe(z) + i dL/dIm(z), which is exactly how the C-ABI
returns gradients of spectra (two real coordinates per stored entry).

    rollout(solver, alpha, W0, R1_0, R2_0, nsteps, ckpt_every) -> (W, R1, R2)
        alpha: real [Ny, Nx] activity field, or [nsteps] uniform activity per
        step.  Forward: nsteps fused IFRK2 steps with a checkpoint every
        ckpt_every steps.  Backward: segment recompute + reverse sweep.
    spectral_loss(mesh, W, kind) -> L
        kind ENSTROPHY: 1/2 mean(omega^2); KINETIC: 1/2 mean(u^2 + v^2).
"""
import torch

from . import ENSTROPHY, KINETIC, adj_loss  # noqa: F401


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alpha, W0, R10, R20, solver, nsteps, ckpt_every):
        field = alpha.dim() == 2
        a = alpha.detach().contiguous()
        solver.set_state(W0.detach().contiguous(), R10.detach().contiguous(), R20.detach().contiguous())
        if field:
            solver.set_activity(a)
            solver.adj_forward(nsteps, ckpt_every, None)
        else:
            if a.numel() != nsteps:
                raise ValueError(f"alpha has {a.numel()} values for {nsteps} steps")
            solver.set_activity(None)
            solver.adj_forward(nsteps, ckpt_every, a)
        W, R1, R2 = solver.get_state()
        ctx.solver, ctx.nsteps, ctx.field = solver, nsteps, field
        ctx.keep = a   # the control must outlive the forward (the library reads it)
        return W, R1, R2

    @staticmethod
    def backward(ctx, gW, gR1, gR2):
        s = ctx.solver
        g = []
        for t in (gW, gR1, gR2):
            g.append(torch.zeros(s.mesh.spec_shape, dtype=torch.complex128, device="cuda")
                     if t is None else t.contiguous())
        want_state = any(ctx.needs_input_grad[1:4])
        ab, at, u = s.adj_backward(*g, want_field=ctx.field, want_sched=not ctx.field, want_state=want_state,
                                   nsteps=ctx.nsteps)
        u = u if u is not None else (None, None, None)
        return (ab if ctx.field else at), u[0], u[1], u[2], None, None, None


def rollout(solver, alpha, W0, R10, R20, nsteps, ckpt_every=0):
    return _Rollout.apply(alpha, W0, R10, R20, solver, int(nsteps), int(ckpt_every))


class _SpectralLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, W, mesh, kind):
        L, Wb = adj_loss(mesh, W.detach().contiguous(), kind)
        ctx.save_for_backward(Wb)
        return torch.tensor(L, dtype=torch.float64, device=W.device)

    @staticmethod
    def backward(ctx, g):
        (Wb,) = ctx.saved_tensors
        return g * Wb, None, None


def spectral_loss(mesh, W, kind=ENSTROPHY):
    return _SpectralLoss.apply(W, mesh, int(kind))
