"""torch.autograd binding of the reverse-mode gradient (salsa_adjoint_device).

    J = precip_volume(ctx, fields, n_p, m_p, E, tstep=1.0, nsteps=20, checkpoint_every=5)
    J.backward()          # n_p[c].grad = dJ/dN0 of every bin, E.grad = dJ/dE

J is the precipitation-bin volume concentration summed over the interior
cells after `nsteps` in-place steps of the SALSA coagulation + condensation
box model (T, p, rs, rsi, dn held; the inputs are not modified).  The
forward call already runs the whole adjoint: the checkpointed forward pass
and the reverse sweep are one C call, whose gradient is kept for backward.
E is the collision-kernel enhancement factor (K = E coagc); the context
fixes its value (1 + charge_coll_enh with lcharge, else 1), and the tensor
passed must hold that value: it is the handle through which dJ/dE reaches
autograd.  dJ/dE excludes the cells flagged SALSA_CELL_GRAD_UNSTABLE, as the
C call's total does.
"""
import torch

from . import capi


def enhancement_factor(ctx):
    """The context's E (mo_salsa_coagulation_kernels.f90:411-416 with
    chargeTime = chargeTimeMax)."""
    p = ctx.params
    return 1.0 + p.charge_coll_enh if p.lcharge else 1.0


class PrecipVolume(torch.autograd.Function):
    @staticmethod
    def forward(fctx, E, sctx, fields, m_p, tstep, nsteps, every, *n_p):
        if abs(float(E) - enhancement_factor(sctx)) > 1e-12 * enhancement_factor(sctx):
            raise ValueError("E must hold the context's enhancement factor %.17g" % enhancement_factor(sctx))
        djdn = [torch.zeros_like(x) for x in n_p]
        J, dJdE = sctx.adjoint_device(fields, list(n_p), m_p, djdn, tstep=tstep, nsteps=nsteps,
                                      checkpoint_every=every)
        fctx.save_for_backward(*djdn)
        fctx.dJdE = dJdE
        return torch.tensor(J, dtype=torch.float64, device=n_p[0].device)

    @staticmethod
    def backward(fctx, gJ):
        djdn = fctx.saved_tensors
        gE = gJ * fctx.dJdE
        return (gE, None, None, None, None, None, None) + tuple(gJ * g for g in djdn)


def precip_volume(ctx, fields, n_p, m_p, E=None, tstep=1.0, nsteps=1, checkpoint_every=5):
    """J as a differentiable function of the number fields n_p (one float64
    tensor per category, the LES layout) and of E (a 0-d float64 tensor; by
    default a new one holding the context's value)."""
    if E is None:
        E = torch.tensor(enhancement_factor(ctx), dtype=torch.float64, device=n_p[0].device)
    return PrecipVolume.apply(E, ctx, fields, m_p, tstep, nsteps, checkpoint_every, *n_p)


__all__ = ["PrecipVolume", "precip_volume", "enhancement_factor", "capi"]
