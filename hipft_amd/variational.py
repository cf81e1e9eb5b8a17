"""Batched gradient assimilation over an observation window.

Each realization of a FlowModel is fitted to its own sequence of observed
maps by minimising the summed window misfit

    J_r = sum_k 1/2 sum_points w_k (f_r(t_k) - obs_k)^2,   t_k = k * steps_per_obs * dt,

with respect to its initial map and log diffusivity / log meridional-flow
amplitude / log rotation scale, using the hand-written adjoint of the
flux-transport step (hipft_amd.adjoint.run, a torch.autograd.Function) and a
torch optimiser.  Realizations are independent, so one batched forward and
one batched adjoint per iteration serve all of them.
"""
import math

import numpy as np
import torch

from . import workloads as W
from .adjoint import FlowModel, run

SYNODIC_DAYS = 27.2753  # Carrington synodic rotation period


class _WeightedMisfit(torch.autograd.Function):
    """J_r = 1/2 sum_points w (x_r - o)^2 per realization; keeps only w (x - o)
    for the backward pass (one field per observation)."""

    @staticmethod
    def forward(ctx, x, o, w):
        d = x - o
        r = d if w is None else w * d
        ctx.save_for_backward(r)
        return 0.5 * (r * d).sum(dim=(1, 2))

    @staticmethod
    def backward(ctx, g):
        (r,) = ctx.saved_tensors
        return g[:, None, None] * r, None, None


def window_misfit(model, f0, eta, v0, omega, obs, weights, steps_per_obs, dt=1.0,
                  checkpoint_every=20, subcycles=1, log=None):
    """Per-realization misfit J_r (tensor [nr]) of the states at the
    observation times against obs[k] ([nr][ntm][nu] each, or [ntm][nu]
    shared by all realizations), weighted by weights[k] (same shape, or None
    for 1).  log(msg), when given, is also called as the adjoint pass reaches
    each observation."""
    x = f0
    J = torch.zeros(f0.shape[0], dtype=torch.float64, device=f0.device)
    for k, o in enumerate(obs):
        x = run(x, eta, v0, omega, model, steps_per_obs, dt=dt, subcycles=subcycles,
                checkpoint_every=checkpoint_every)
        if log is not None and x.requires_grad:
            x.register_hook(lambda g, k=k: log(f"adjoint pass at observation {k}"))
        w = None if weights is None else weights[k]
        J = J + _WeightedMisfit.apply(x, o, w)
    return J


def fit(model, f0, eta0, v0_0, omega0, obs, weights, steps_per_obs, iters=50, dt=1.0,
        fit_map=True, fit_eta=True, fit_v0=True, fit_omega=False, lr_log=0.05, lr_map=1.0,
        checkpoint_every=20, subcycles=1, log=None):
    """Adam on the chosen parameters; returns the fitted values and the loss
    history (sum over realizations)."""
    dev = f0.device
    to = lambda a: torch.as_tensor(a, dtype=torch.float64).clone()
    le = torch.log(to(eta0)).requires_grad_(fit_eta)
    lv = torch.log(to(v0_0)).requires_grad_(fit_v0)
    lo = torch.log(to(omega0)).requires_grad_(fit_omega)
    x0 = f0.detach().clone().to(dev).requires_grad_(fit_map)
    groups = []
    logs = [p for p, on in ((le, fit_eta), (lv, fit_v0), (lo, fit_omega)) if on]
    if logs:
        groups.append({"params": logs, "lr": lr_log})
    if fit_map:
        groups.append({"params": [x0], "lr": lr_map})
    opt = torch.optim.Adam(groups)
    history = []
    for it in range(iters):
        opt.zero_grad()
        J = window_misfit(model, x0, torch.exp(le), torch.exp(lv), torch.exp(lo), obs, weights,
                          steps_per_obs, dt=dt, checkpoint_every=checkpoint_every,
                          subcycles=subcycles)
        total = J.sum()
        total.backward()
        opt.step()
        history.append(float(total))
        if log:
            log(f"iteration {it}: misfit {history[-1]:.6e}")
    return dict(eta=torch.exp(le).detach(), v0=torch.exp(lv).detach(),
                omega=torch.exp(lo).detach(), f0=x0.detach(), loss=history)


def synthetic_window(ntm, npm1, nr, ndays, seed=350, eta_truth=350.0, v0_truth=15.0,
                     noise_g=5.0, mu_power=4.0, mu_limit=0.5, nbip=200, clustered=True,
                     steps_per_day=24, autotune=True):
    """A batched gradient-assimilation problem: one seeded truth run (eta_truth,
    meridional amplitude v0_truth, unit rotation scale) observed every
    steps_per_day steps of 1 h (once a day by default), ndays times, as
    truth + N(0, noise_g) Gauss, each observation weighted by
    mu^mu_power where mu > mu_limit (the disk within 60 degrees of centre for
    0.5) and 0 elsewhere, mu = sin(theta) cos(phi - phi_c) for an Earth
    observer whose Carrington longitude phi_c falls by 360 degrees per
    synodic rotation; the nr realizations start from guesses eta ~
    log-uniform(100, 1000) km^2/s, v0 ~ uniform(10, 20) m/s and the truth's
    initial map plus N(0, noise_g) noise.  Observations and weights are
    [ntm][npm-2] tensors shared by all realizations."""
    t = W.clustered_theta(ntm, 2.0) if clustered else W.uniform_theta(ntm)
    p = W.uniform_phi(npm1)
    rng = np.random.default_rng(seed)
    vt1 = W.meridional_profile(t, 1.0)[None]
    dr1 = np.array([W.dr_coefficients(1.0)])
    truth = FlowModel(t, p, 1, vt_unit=vt1, dr_unit=dr1, autotune=autotune)
    truth.ft.set_state(W.bipole_map(t, p, nbip, seed + 1), layout="pt")
    truth.bind_stream()
    x = truth.get()
    x_init = x[0].clone()
    one = lambda v: torch.tensor([v], dtype=torch.float64)
    obs, weights = [], []
    with torch.no_grad():
        for k in range(1, ndays + 1):
            x = run(x, one(eta_truth), one(v0_truth), one(1.0), truth, steps_per_day)
            noise = torch.as_tensor(rng.normal(0.0, noise_g, tuple(x.shape[1:])), device="cuda")
            obs.append(x[0] + noise)
            days = k * steps_per_day / 24.0  # dt = 1 h
            mu = W.mu_map(t, p, -days * 360.0 / SYNODIC_DAYS)[:, 1:]  # unique columns
            w = np.where(mu > mu_limit, np.power(np.abs(mu), mu_power), 0.0)
            weights.append(torch.as_tensor(w, device="cuda"))
    truth.ft.close()
    vt_unit = np.repeat(vt1, nr, axis=0)
    dr_unit = np.repeat(dr1, nr, axis=0)
    model = FlowModel(t, p, nr, vt_unit=vt_unit, dr_unit=dr_unit, autotune=autotune)
    model.bind_stream()
    f0 = x_init[None] + torch.as_tensor(rng.normal(0.0, noise_g, (nr,) + tuple(x_init.shape)),
                                        device="cuda")
    eta0 = torch.tensor(np.exp(rng.uniform(math.log(100.0), math.log(1000.0), nr)),
                        dtype=torch.float64)
    v0_0 = torch.tensor(rng.uniform(10.0, 20.0, nr), dtype=torch.float64)
    omega0 = torch.ones(nr, dtype=torch.float64)
    return dict(model=model, t=t, p=p, f0=f0, eta0=eta0, v0_0=v0_0, omega0=omega0, obs=obs,
                weights=weights, steps_per_obs=steps_per_day, truth_f0=x_init,
                eta_truth=eta_truth, v0_truth=v0_truth)
