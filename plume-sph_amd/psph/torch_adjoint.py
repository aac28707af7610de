"""torch.autograd.Function over the plume adjoint (psph_adjoint_run).

``plume_height(vent_v, vent_T, mass_frac, dv0, ...)`` builds the weak-plume
state for the given vent parameters (psph_gen_plume), adds the per-particle
initial-velocity perturbation ``dv0`` (n, 3) to the velocity v, runs
``nsteps`` steps on the GPU and returns the diagnostic H (mass-weighted mean
height of erupted material).  Its backward is the hand-written adjoint:
dH/d(vent_v, vent_T, mass_frac) and dH/d dv0 = dH/dv of the starting state.
The forward computes the gradients eagerly (one reverse sweep per call);
torch only routes them.  A ``dv0`` on the GPU is applied in place through the
library's zero-copy device view (psph_device_view), not through host memory.

The smoothed velocity sv is the velocity filter's state and stays as
generated.  (In a fluid at rest H is not differentiable in sv: the heat
conduction term of a pair is proportional to |v_ij| and the artificial
viscosity switches on at v_ij . r_ij = 0.)

mass_frac is the erupted solid mass fraction, 1 - ng0.
"""
from __future__ import annotations

import numpy as np
import torch

import psph


def build_state(cfg: psph.Config, met, seed: int, jitter: float, vent_v: float, vent_T: float,
                mass_frac: float, dv0: np.ndarray | None):
    """Config with the given vent and the generated particles (v += dv0)."""
    c = cfg.copy()
    c.vent_v, c.vent_T, c.ng0 = float(vent_v), float(vent_T), 1.0 - float(mass_frac)
    psph.check(psph.lib().psph_config_vent(psph.C.byref(c)))
    P = psph.gen_plume(c, met, seed, jitter)
    if dv0 is not None:
        dv0 = np.asarray(dv0, np.float64).reshape(len(P), 3)
        for k, f in enumerate(("vx", "vy", "vz")):
            P[f][:] += dv0[:, k]
    return c, P


class PlumeHeight(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vent_v, vent_T, mass_frac, dv0, cfg, met, seed, jitter, nsteps, ckpt_every, extra_capacity):
        on_device = dv0.is_cuda
        c, P = build_state(cfg, met, seed, jitter, vent_v.item(), vent_T.item(), mass_frac.item(),
                           None if on_device else dv0.detach().cpu().numpy())
        with psph.Solver(c, len(P) + extra_capacity) as S:
            S.load(P)
            if on_device:
                # dv0 stays on the device: added in place through the zero-copy
                # view, whose order right after the load is P's
                if dv0.shape != (len(P), 3):
                    raise ValueError(f"dv0 has shape {tuple(dv0.shape)}, the state {len(P)} particles")
                V = S.device_view()
                d = dv0.detach().to(torch.float64)
                for k, f in enumerate(("vx", "vy", "vz")):
                    V[f].add_(d[:, k])
                torch.cuda.synchronize()
                del V
            r = S.adjoint_run(nsteps, ckpt_every, grads=True)
        if r["n_particles"] != len(P):
            raise psph.PsphError(-6, f"adjoint returned {r['n_particles']} gradients for {len(P)} particles")
        gdv = torch.from_numpy(r["grad_v"]).to(dv0.dtype).to(dv0.device)
        ctx.save_for_backward(torch.tensor(r["dH_dvent_v"], dtype=torch.float64),
                              torch.tensor(r["dH_dvent_T"], dtype=torch.float64),
                              torch.tensor(r["dH_dmass_frac"], dtype=torch.float64), gdv)
        return torch.tensor(r["H"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gH):
        gv, gT, gm, gdv = ctx.saved_tensors
        return (gH * gv, gH * gT, gH * gm, gH.to(gdv.dtype) * gdv,
                None, None, None, None, None, None, None)


def plume_height(vent_v, vent_T, mass_frac, dv0, *, cfg, met, seed, jitter, nsteps, ckpt_every,
                 extra_capacity=200000):
    """H after nsteps steps; differentiable in the four tensor arguments."""
    return PlumeHeight.apply(vent_v, vent_T, mass_frac, dv0, cfg, met, seed, jitter, nsteps, ckpt_every,
                             extra_capacity)
