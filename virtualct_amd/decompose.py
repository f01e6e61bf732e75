"""Dual-energy material decomposition by projected gradient descent.

Unknowns: elemental partial densities rho[b, m, y, x] (g/cm^3, float32) of a
batch of phantoms.  Data: noisy normalised sinograms t[b, s] of several
spectra (120 kV and 6 MV in the benchmark).  Loss per phantom

    l_b = sum_s sum_rays (ln p_bs - ln t_bs)^2,   p_bs = I_bs(rho_b) / I_max,bs

with I the polyenergetic forward model (MBFPA) and I_max held constant (it is
sum_E pdf whenever a ray misses the object).  The gradient comes from the
hand-written adjoint.  Updates are projected Barzilai-Borwein steps,
rho <- max(0, rho - alpha_b * g_b), alpha_b from the last two iterates.

Two drivers share the mathematics:
  * ``PlainLoop``: per phantom, the plain forward of all spectra from one
    trace and one plain adjoint for their combined gradient (what a user of
    the single-phantom API writes);
  * ``BatchedSolver`` (virtualct_amd.batch): every phantom and spectrum in
    one launch per stage over the cached system matrix, captured in a HIP
    graph.
"""
from __future__ import annotations

import torch

from . import ops


def log_residual(I: torch.Tensor, imax: torch.Tensor, target: torch.Tensor):
    """r = ln(I/imax) - ln t and d l / d I = 2 r / I (elementwise)."""
    p = I / imax
    r = torch.log(p) - torch.log(target)
    return r, 2.0 * r / I


def reprojection_error(plan, dens: torch.Tensor, target: torch.Tensor, att: torch.Tensor, pdf: torch.Tensor):
    """Relative RMS of ln p - ln t over all rays and spectra for one phantom."""
    with torch.no_grad():
        I = ops.mbfpa_intensity(dens, plan, att, pdf)
        imax = I.amax(dim=(1, 2), keepdim=True)
        r, _ = log_residual(I, imax, target)
        return (r.square().mean() / torch.log(target).square().mean()).sqrt().item()


class PlainLoop:
    """Per-phantom projected-gradient loop over the plain API: one
    ``mbfpa_intensity`` call (one trace, every spectrum) and one backward
    (one adjoint) per phantom and iteration."""

    def __init__(self, plan, targets, att, pdf, init):
        self.plan = plan
        self.t = targets          # [B,S,V,W]
        self.att = att            # [S,M,NE]
        self.pdf = pdf            # [S,NE]
        self.x = init.clone()     # [B,M,N,N]
        self.B, self.S = targets.shape[:2]
        self.M = init.shape[1]
        self.prev_x = None
        self.prev_g = None
        self.alpha = torch.full((self.B,), 1e-4, device=init.device)

    def gradient(self, b):
        x = self.x[b].detach().requires_grad_(True)
        I = ops.mbfpa_intensity(x, self.plan, self.att, self.pdf)  # [S,V,W]
        imax = I.detach().amax(dim=(1, 2), keepdim=True)
        r, dI = log_residual(I.detach(), imax, self.t[b])
        I.backward(dI)
        return x.grad, r.square().sum()

    def step(self):
        losses = []
        new_x = torch.empty_like(self.x)
        grads = torch.empty_like(self.x)
        for b in range(self.B):
            g, loss = self.gradient(b)
            grads[b] = g
            losses.append(loss)
        if self.prev_x is not None:
            s = (self.x - self.prev_x).flatten(1)
            y = (grads - self.prev_g).flatten(1)
            sy = (s * y).sum(1)
            ss = (s * s).sum(1)
            bb = torch.where(sy > 0, ss / sy.clamp_min(1e-30), self.alpha * 2)
            self.alpha = bb.clamp(1e-8, 1e2)
        new_x = (self.x - self.alpha.view(-1, 1, 1, 1) * grads).clamp_min(0.0)
        self.prev_x, self.prev_g = self.x, grads
        self.x = new_x
        return torch.stack(losses)


def water_init(target_fbp_mu: torch.Tensor, M: int, mu_water: float) -> torch.Tensor:
    """Water-equivalent start: H and O of water scaled by mu_FBP / mu_water on
    the support mu_FBP > 0.2 mu_water.  target_fbp_mu: [B,N,N] in 1/cm."""
    scale = (target_fbp_mu / mu_water).clamp(0.0, 3.0) * (target_fbp_mu > 0.2 * mu_water)
    x = torch.zeros((target_fbp_mu.shape[0], M) + tuple(target_fbp_mu.shape[1:]), device=target_fbp_mu.device)
    x[:, 0] = 0.1119 * scale
    x[:, 3] = 0.8881 * scale
    return x
