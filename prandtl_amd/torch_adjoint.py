"""Thin torch binding of the discrete adjoint (tests and the adjoint benchmark).

RKSolve.apply(u0, op, dt, nsteps, ckpt_every[, scheme[, t0]]) advances u0 by nsteps LSRK4(5) (or
SSP-RK3) steps from time t0 through the C ABI (prandtl_rk_run_tape) and, on backward, returns
dJ/du0 = (du_T/du0)^T dJ/du_T from prandtl_rk_adjoint: the hand-written operator adjoint and the
reverse RK stages, with the stage inputs recomputed segment by segment from the checkpoints.
The handle keeps one tape: backward checks that it is still the forward's (check_tape).
"""
from __future__ import annotations

import torch


class TapeMismatch(RuntimeError):
    """The operator's last taped run is not the one a backward pass differentiates."""


def check_tape(op, tape_id):
    """Raise TapeMismatch unless the handle's last taped run is `tape_id` (prandtl_rk_adjoint sweeps
    the handle's last tape: a later taped run on the same operator replaces it)."""
    cur = op.tape_id()
    if tape_id == 0 or cur != tape_id:
        raise TapeMismatch(f"the operator's last taped run is {cur}, this backward needs run {tape_id}: "
                           f"another taped run on the same operator replaced it (run the backward first, or "
                           f"use one operator per trajectory)")


class RKSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, op, dt, nsteps, ckpt_every, scheme=1, t0=0.0):
        u = u0.detach().clone().contiguous()
        op.rk_run_tape(u, float(t0), dt, nsteps, ckpt_every, scheme=scheme)
        ctx.op = op
        ctx.tape_id = op.tape_id()
        return u

    @staticmethod
    def backward(ctx, ubar):
        check_tape(ctx.op, ctx.tape_id)
        g = ubar.detach().clone().contiguous()
        ctx.op.rk_adjoint(g)
        return g, None, None, None, None, None, None


def central_fd(f, x, h):
    """Central differences of a scalar function f of the scalar x: returns (D4(h), D4(2h), D6)
    with D4(h) = (8 (f(x+h) - f(x-h)) - (f(x+2h) - f(x-2h))) / (12 h) (fourth order) and the
    sixth-order Richardson combination D6 = (16 D4(h) - D4(2h)) / 15."""
    fp = {k: f(x + k * h) for k in (-4, -2, -1, 1, 2, 4)}
    d4h = (8.0 * (fp[1] - fp[-1]) - (fp[2] - fp[-2])) / (12.0 * h)
    d42h = (8.0 * (fp[2] - fp[-2]) - (fp[4] - fp[-4])) / (24.0 * h)
    return d4h, d42h, (16.0 * d4h - d42h) / 15.0


def kinetic_energy(u, W):
    """J = sum_nodes W |rho v|^2 / (2 rho): the volume-integrated kinetic energy (GLL quadrature)."""
    m2 = sum(u[1 + d] ** 2 for d in range(u.shape[0] - 2))
    return (W * 0.5 * m2 / u[0]).sum()


def tgv_modes_state(mesh, mach, amps, modes, gamma=1.4):
    """Differentiable (in mach and amps) TGV + Fourier-mode initial state, torch float64."""
    import numpy as np

    ks, dirs, phases = (torch.as_tensor(np.asarray(a), dtype=torch.float64, device=amps.device) for a in modes)
    X = torch.as_tensor(mesh.node_coords(), dtype=torch.float64, device=amps.device)
    x, y, z = X
    rho = torch.ones_like(x)
    u = torch.sin(x) * torch.cos(y) * torch.cos(z)
    v = -torch.cos(x) * torch.sin(y) * torch.cos(z)
    w = torch.zeros_like(x)
    arg = ks[:, 0, None, None] * x + ks[:, 1, None, None] * y + ks[:, 2, None, None] * z + phases[:, None, None]
    s = amps[:, None, None] * torch.sin(arg)
    u = u + (dirs[:, 0, None, None] * s).sum(0)
    v = v + (dirs[:, 1, None, None] * s).sum(0)
    w = w + (dirs[:, 2, None, None] * s).sum(0)
    p = 1.0 / (gamma * mach**2) + (1.0 / 16.0) * (torch.cos(2 * x) + torch.cos(2 * y)) * (torch.cos(2 * z) + 2.0)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return torch.stack([rho, rho * u, rho * v, rho * w, E])
