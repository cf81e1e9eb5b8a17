"""Differentiable flux transport: a torch.autograd.Function over libhft.

``FluxTransportRun.apply(f0, eta, v0, omega, model, nsteps, ...)`` advances
the batch of maps ``f0`` (a CUDA float64 tensor [nr][ntm][npm-2], the unique
phi columns) by ``nsteps`` flux-transport steps on the GPU and returns the
final maps.  Its backward is the hand-written discrete adjoint of libhft
(``hft_step_adjoint``) with step checkpointing: the forward keeps the state
every ``checkpoint_every`` steps; the backward recomputes each segment from
its checkpoint (one forward pass keeping the stage data, when it fits in
the library's adjoint memory budget) and runs the step adjoints in reverse.

Per-realization parameters (tensors [nr], float64):
  eta    -- diffusivity (km^2/s),
  v0     -- amplitude multiplying the model's meridional-flow profile,
  omega  -- scale multiplying the model's differential-rotation coefficients.
Gradients are exact for the discrete step with its discrete choices held
fixed (RKG2 stage counts, sub-cycle time steps, upwind directions).
"""
import numpy as np
import torch

from .core import FluxTransport


class FlowModel:
    """A FluxTransport context whose flows and diffusivity are functions of
    (eta, v0, omega) per realization:
      vt = v0_r * vt_unit_r (theta half mesh, code units),
      (p0, p2, p4) = omega_r * dr_unit_r (m/s),
    with no other flow terms, so that d/dlog(v0) and d/dlog(omega) are the
    library's theta- and phi-velocity scale sensitivities."""

    def __init__(self, t, p, nr, vt_unit=None, dr_unit=None, mode="independent", autotune=True):
        self.ft = FluxTransport(t, p, nr, mode=mode)
        self.autotune = autotune
        self.nr = nr
        self.nu = self.ft.npm - 2
        self.ntm = self.ft.ntm
        self.vt_unit = None if vt_unit is None else np.asarray(vt_unit, dtype=np.float64)
        self.dr_unit = None if dr_unit is None else np.asarray(dr_unit, dtype=np.float64)
        self._params = None

    def apply_params(self, eta, v0, omega):
        eta, v0, omega = (np.asarray(a, dtype=np.float64).reshape(self.nr) for a in (eta, v0, omega))
        key = (eta.tobytes(), v0.tobytes(), omega.tobytes())
        if key == self._params:
            return
        ft = self.ft
        ft.set_diffusion(eta)
        if self.vt_unit is not None or self.dr_unit is not None:
            if self.vt_unit is not None:
                ft.set_flow_profiles(v0[:, None] * self.vt_unit, None)
            if self.dr_unit is not None:
                d = omega[:, None] * self.dr_unit
                ft.set_flow_analytic(p0=d[:, 0], p2=d[:, 1], p4=d[:, 2])
            elif self.vt_unit is not None:
                ft.set_flow_analytic()
        if self.autotune and self._params is None:
            from .tune import autotune
            # launch shapes for this grid and batch (cached; a batch size not
            # measured yet borrows the nearest one's)
            autotune(ft, borrow=True)
        self._params = key

    def bind_stream(self):
        self.ft.set_stream(torch.cuda.current_stream().cuda_stream)

    def put(self, x):
        self.ft.set_state_device(x.data_ptr(), self.nu)

    def get(self):
        out = torch.empty((self.nr, self.ntm, self.nu), dtype=torch.float64, device="cuda")
        self.ft.get_state_device(out.data_ptr(), self.nu)
        return out


class FluxTransportRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, eta, v0, omega, model, nsteps, dt=1.0, strang=True, subcycles=1,
                checkpoint_every=20, frozen_weno=False):
        if not f0.is_cuda or f0.dtype != torch.float64:
            raise TypeError("f0 must be a CUDA float64 tensor")
        model.bind_stream()
        model.apply_params(eta.detach().cpu().numpy(), v0.detach().cpu().numpy(),
                           omega.detach().cpu().numpy())
        x = f0.detach().contiguous()
        model.put(x)
        ckpts = {}
        for n in range(nsteps):
            if n % checkpoint_every == 0:
                ckpts[n] = x if n == 0 else model.get()
            model.ft.flux_transport_step(dt, strang=strang, subcycles=subcycles)
        out = model.get()
        ctx.model = model
        ctx.ckpts = ckpts
        ctx.cfg = (nsteps, dt, strang, subcycles, frozen_weno)
        ctx.save_for_backward(eta, v0, omega)
        return out

    @staticmethod
    def backward(ctx, g):
        model = ctx.model
        nsteps, dt, strang, subcycles, frozen = ctx.cfg
        eta, v0, omega = ctx.saved_tensors
        model.bind_stream()
        model.apply_params(eta.detach().cpu().numpy(), v0.detach().cpu().numpy(),
                           omega.detach().cpu().numpy())
        lam = g.detach().contiguous().clone()
        grads = np.zeros((model.nr, 3))
        starts = sorted(ctx.ckpts)
        for si in reversed(range(len(starts))):
            n0 = starts[si]
            n1 = starts[si + 1] if si + 1 < len(starts) else nsteps
            model.put(ctx.ckpts[n0])
            grads += model.ft.steps_adjoint(n1 - n0, dt, lam.data_ptr(), model.nu, strang=strang,
                                            subcycles=subcycles, frozen_weno=frozen)
        dev = eta.device
        t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev)
        ge = t(grads[:, 0]) / eta.detach()
        gv = t(grads[:, 1]) / v0.detach()
        go = t(grads[:, 2]) / omega.detach()
        return lam, ge, gv, go, None, None, None, None, None, None, None


def run(f0, eta, v0, omega, model, nsteps, dt=1.0, strang=True, subcycles=1, checkpoint_every=20,
        frozen_weno=False):
    """Differentiable nsteps-step flux transport of the batch f0."""
    return FluxTransportRun.apply(f0, eta, v0, omega, model, nsteps, dt, strang, subcycles,
                                  checkpoint_every, frozen_weno)
