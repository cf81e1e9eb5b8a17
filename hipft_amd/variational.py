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
import torch

from .adjoint import run


def window_misfit(model, f0, eta, v0, omega, obs, weights, steps_per_obs, dt=1.0,
                  checkpoint_every=20, subcycles=1):
    """Per-realization misfit J_r (tensor [nr]) of the states at the
    observation times against obs[k] ([nr][ntm][nu] each), weighted by
    weights[k] (same shape, or None for 1)."""
    x = f0
    J = torch.zeros(f0.shape[0], dtype=torch.float64, device=f0.device)
    for k, o in enumerate(obs):
        x = run(x, eta, v0, omega, model, steps_per_obs, dt=dt, subcycles=subcycles,
                checkpoint_every=checkpoint_every)
        d = x - o
        w = None if weights is None else weights[k]
        e = d * d if w is None else w * d * d
        J = J + 0.5 * e.sum(dim=(1, 2))
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
