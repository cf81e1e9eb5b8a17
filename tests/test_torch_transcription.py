"""The differentiable PyTorch transcription of the step (oracle/torch_step.py,
used to check the GPU adjoint) against the C oracle, on the CPU."""
import numpy as np
import pytest
import torch

from hipft_amd import workloads as W
from oracle.oracle import Oracle, sts_rkg2
from oracle.torch_step import TorchStep


def setup(ntm=25, npm1=49, nr=2, seed=3, eta=(300.0, 300.0)):
    from hipft_amd.core import DIFF_KM2_S_TO_RS2_HR
    t, p = W.clustered_theta(ntm, 2.0), W.uniform_phi(npm1)
    o = Oracle(t, p, nr)
    maps = np.stack([W.bipole_map(t, p, 6, seed + r) for r in range(nr)])
    o.ingest_pt(maps)
    vt = np.stack([W.meridional_profile(t, 12.0 + 4.0 * r) for r in range(nr)])
    o.flow_reset()
    o.flow_set(np.ascontiguousarray(np.broadcast_to(vt[:, None, :], (nr, o.npm, o.nt))),
               np.zeros((nr, o.np_, o.ntm)))
    p0, p2, p4 = (np.array([c * (1.0 + 0.05 * r) for r in range(nr)]) for c in W.dr_coefficients())
    o.flow_add_dr(p0, p2, p4)
    o.set_lf_alpha()
    o.load_diffusion(list(eta))
    D = np.array(eta) * DIFF_KM2_S_TO_RS2_HR
    return o, TorchStep(o, D=D)


def unique(o, a):
    """(nr, npm, ntm) reference layout -> (nr, ntm, nu) unique columns."""
    return torch.as_tensor(np.ascontiguousarray(a[:, 1:o.npm - 1, :].transpose(0, 2, 1)))


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def params(nr):
    z = torch.zeros(nr, dtype=torch.float64)
    return z.clone(), z + 1.0, z + 1.0


def test_operators_match_oracle():
    o, ts = setup()
    x = unique(o, o.get_f())
    logD, st, sp = params(o.nr)
    assert rel(ts.W(x, st, sp), unique(o, o.advection_operator(o.get_f()))) < 1e-13
    assert rel(ts.L(x, logD), unique(o, o.diffusion_operator(o.get_f()))) < 1e-13


def test_step_matches_oracle():
    o, ts = setup()
    x = unique(o, o.get_f())
    s = sts_rkg2(1.0, o.dt_euler)[0]
    logD, st, sp = params(o.nr)
    y = ts.step(x, logD, st, sp, 1.0, strang=True, subcycles=1, s_r=[s] * o.nr)
    o.advection_step(0.5)
    o.diffusion_step(1.0, subcycles=1)
    o.advection_step(0.5)
    assert rel(y, unique(o, o.get_f())) < 1e-12


def test_autograd_matches_central_differences():
    """Autograd of the transcription equals central differences of itself
    (no non-differentiable operation on the path)."""
    o, ts = setup(ntm=17, npm1=33)
    x0 = unique(o, o.get_f())
    rng = np.random.default_rng(0)
    tgt = torch.as_tensor(rng.standard_normal(x0.shape))
    s_r = ts.stage_counts(1.0)

    def J(x, logD, st, sp):
        y = ts.step(x, logD, st, sp, 1.0, strang=True, subcycles=1, s_r=s_r)
        return 0.5 * ((y - tgt) ** 2).sum()

    logD, st, sp = params(o.nr)
    x = x0.clone().requires_grad_(True)
    for v in (logD, st, sp):
        v.requires_grad_(True)
    J(x, logD, st, sp).backward()
    dx = torch.as_tensor(rng.standard_normal(x0.shape))
    h = 1e-4  # J ~ 4e4: smaller steps are dominated by rounding of J
    with torch.no_grad():
        fd = (J(x0 + h * dx, logD, st, sp) - J(x0 - h * dx, logD, st, sp)) / (2 * h)
        ad = (x.grad * dx).sum()
        assert abs(float(fd - ad)) <= 1e-7 * abs(float(ad))
        for name, v in (("logD", logD), ("st", st), ("sp", sp)):
            e = torch.zeros_like(v)
            e[1] = 1.0
            args = dict(logD=logD, st=st, sp=sp)
            ap = dict(args)
            am = dict(args)
            ap[name] = v + h * e
            am[name] = v - h * e
            fd = (J(x0, **ap) - J(x0, **am)) / (2 * h)
            assert abs(float(fd - v.grad[1])) <= 1e-7 * abs(float(v.grad[1])) + 1e-9, name
