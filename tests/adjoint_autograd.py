"""Test-only torch.autograd.Function over the library's discrete adjoint:
the final (Y, T) of one CFD chemistry step as a differentiable function of
the initial (Y, T) of every cell and of the mechanism's rate parameters
(ln|A|, beta, Ta per rate expression).  Forward runs gfc_adjoint_forward,
backward gfc_adjoint_backward (the gradient of the discrete map with the
forward's accepted step sizes held fixed)."""
from __future__ import annotations

import torch


class ChemistryStep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state0, theta, ev, deltaT, dtChem0, p):
        """state0 [nS+1, nCells] (cuda f64), theta [nR, 6] (f64), p [nCells]."""
        ev.set_rate_parameters(theta.detach().cpu().numpy())
        Y0 = torch.cat([state0.detach(), p.view(1, -1)], 0).contiguous()
        Yn = torch.empty_like(Y0)
        dt = dtChem0.clone()
        ev.adjoint_forward(deltaT, Y0, dt, Ynew=Yn)
        ctx.ev = ev
        ctx.theta_device = theta.device
        return Yn[:-1].clone()

    @staticmethod
    def backward(ctx, gout):
        ev = ctx.ev
        Ybar = gout.contiguous()
        Y0bar = torch.empty_like(Ybar)
        pbar = torch.empty((ev.n_params, Ybar.shape[1]), dtype=torch.float64,
                           device=Ybar.device)
        ev.adjoint_backward(Ybar, Y0bar, pbar)
        gtheta = pbar.sum(1).view(-1, 6).to(ctx.theta_device)
        return Y0bar, gtheta, None, None, None, None


def replay(ev, state0, theta, p):
    """The map ChemistryStep differentiates: the last recorded accepted
    steps re-integrated from state0 with parameters theta."""
    ev.set_rate_parameters(theta.detach().cpu().numpy())
    Y0 = torch.cat([state0.detach(), p.view(1, -1)], 0).contiguous()
    Yn = torch.empty_like(Y0)
    ev.adjoint_replay(Y0, Yn)
    return Yn[:-1]
