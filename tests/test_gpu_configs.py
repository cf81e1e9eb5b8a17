"""GPU parity at the shapes of the five benchmark configurations.

1. 181x360 uniform, one realization, 6 bipoles, flows, eta = 500, Strang,
   dt = 1 h: two days against the oracle, bit for bit.
2. 512x1024 clustered, one realization, 200 bipoles: one day, then a
   mu^4-weighted assimilation of the state plus 5 G noise within 60 degrees
   of disk centre; bit for bit before the assimilation, 1e-13 after (pow).
3. The 64-realization benchmark batch (independent stage counts) with the
   tuned launch shapes: two steps, every realization against a single-
   realization oracle run, bit for bit.
4. The adjoint configuration (256x512, 16 realizations, 200 steps,
   checkpoints every 20): the gradient of the final misfit against central
   differences of the forward HIP run on 32 map directions and the three
   scalars of two realizations, agreement 1e-6.
5. The gradient-assimilation shape (512x1024, 128 realizations): a two-step
   window's eta gradient against central differences with the device memory
   held, and a three-day mu^4-weighted noisy window fitted for three Adam
   iterations on eta, v0 and the map, with its gradients checked.
"""
import math
import os
import sys

import numpy as np
import pytest
import torch

from hipft_amd import workloads as W
from hipft_amd.core import FluxTransport
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def assert_same(a, b, what, rtol=0.0):
    if rtol == 0.0:
        bad = np.argwhere(a != b)
        assert bad.size == 0, f"{what}: {len(bad)} entries differ, max |diff| " \
                              f"{np.abs(a - b).max():.3e}"
    else:
        err = np.abs(a - b).max() / np.abs(b).max()
        assert err <= rtol, f"{what}: max rel error {err:.3e} > {rtol:.1e}"


def oracle_with(t, p, maps, vt, dr, eta):
    """Single- or multi-realization oracle with 1-D meridional profiles vt
    [nr][nt], DR coefficients dr [nr][3] and diffusivities eta [nr]."""
    nr = len(eta)
    o = Oracle(t, p, nr)
    o.ingest_pt(maps)
    o.flow_reset()
    o.flow_set(np.ascontiguousarray(np.broadcast_to(vt[:, None, :], (nr, o.npm, o.nt))),
               np.zeros((nr, o.np_, o.ntm)))
    o.flow_add_dr(dr[:, 0], dr[:, 1], dr[:, 2])
    o.set_lf_alpha()
    o.load_diffusion(list(eta))
    return o


def gpu_with(t, p, maps, vt, dr, eta, mode):
    g = FluxTransport(t, p, len(eta), mode=mode)
    g.set_state(maps, layout="pt")
    g.set_flow_profiles(vt, None)
    g.set_flow_analytic(p0=dr[:, 0], p2=dr[:, 1], p4=dr[:, 2])
    g.set_diffusion(eta)
    return g


def oracle_step(o, dt=1.0):
    o.advection_step(0.5 * dt)
    o.diffusion_step(dt, 0)
    o.advection_step(0.5 * dt)


def test_config1_plain_two_days(gpu_available):
    t, p = W.uniform_theta(181), W.uniform_phi(360)
    maps = W.bipole_map(t, p, 6, 1, fixed_sigma=3.0, fixed_b=100.0)[None]
    vt = W.meridional_profile(t, 15.0)[None]
    dr = np.array([W.dr_coefficients()])
    eta = np.array([500.0])
    o = oracle_with(t, p, maps, vt, dr, eta)
    g = gpu_with(t, p, maps, vt, dr, eta, "coupled")
    for _ in range(48):
        oracle_step(o)
        g.flux_transport_step(1.0, strang=True, subcycles=0)
    assert_same(g.get_state("tp"), o.get_f(), "181x360, 48 steps")
    h_g, h_o = g.analysis()[0], o.analysis()[0]
    assert np.allclose(h_g, h_o, rtol=1e-12, atol=1e-12 * abs(h_o[3]))
    g.close()


def test_config2_day_and_assimilation(gpu_available):
    t, p = W.clustered_theta(512, 2.0), W.uniform_phi(1024)
    maps = W.bipole_map(t, p, 200, 2)[None]
    vt = W.meridional_profile(t, 15.0)[None]
    dr = np.array([W.dr_coefficients()])
    eta = np.array([500.0])
    o = oracle_with(t, p, maps, vt, dr, eta)
    g = gpu_with(t, p, maps, vt, dr, eta, "coupled")
    for _ in range(24):
        oracle_step(o)
        g.flux_transport_step(1.0, strang=True, subcycles=0)
    assert_same(g.get_state("tp"), o.get_f(), "512x1024, 24 steps")
    rng = np.random.default_rng(3)
    obs = g.get_state("pt")[0] + rng.normal(0.0, 5.0, (len(t), len(p)))
    d3 = np.stack([obs, np.zeros_like(obs), W.mu_map(t, p, 40.0)])
    kw = dict(custom_from_mu=True, mu_power=4.0, mu_limit=0.5)
    o.assimilate(d3[None], **kw)
    g.assimilate(d3, **kw)
    assert_same(g.get_state("tp"), o.get_f(), "assimilation at 512x1024", 1e-13)
    g.close()


def test_config3_batch_steps_vs_single_runs(gpu_available):
    """All 64 realizations of the benchmark batch with the launch shapes
    bench.py uses (tuning cache: fused multi-stage RKG2 launches, banded
    WENO3), two Strang steps with PTL sub-cycling: each realization equals a
    single-realization oracle run with its own parameters, bit for bit."""
    sys.path.insert(0, ROOT)
    import bench
    from hipft_amd import tune
    wl = bench.make_workload(64)
    g = bench.setup_model(wl)
    cfg = tune.autotune(g)
    print(f"tuned launch shapes: {cfg}", flush=True)
    for _ in range(2):
        g.flux_transport_step(1.0, strang=True, subcycles=0)
    f = g.get_state("tp")
    s, cyc = g.diffusion_stats()
    g.close()
    bad = []
    for r in range(64):
        o = oracle_with(wl["t"], wl["p"], wl["maps"][r:r + 1], wl["vt"][r:r + 1],
                        wl["dr"][r:r + 1], wl["eta"][r:r + 1])
        oracle_step(o)
        oracle_step(o)
        if not np.array_equal(f[r:r + 1], o.get_f()):
            bad.append((r, float(np.abs(f[r:r + 1] - o.get_f()).max())))
    print(f"64 realizations, stage counts {s.min()}..{s.max()}, cycles {cyc.min()}..{cyc.max()}",
          flush=True)
    assert not bad, f"realizations differing from their oracle runs: {bad[:5]}"
    assert s.max() > 2 * s.min()  # the batch's stage counts differ


def test_config3_tuned_batch_conserves_signed_flux(gpu_available):
    """The benchmark batch exactly as bench.py runs it (64 realizations at
    512x1024, launch shapes from the tuning cache, PTL sub-cycling, Strang
    steps): the signed flux of every realization, weighted as the flux-form
    operators conserve it (interior cells sin(t) dt dp, each pole value
    pi sin(th) dth / 2), stays at its initial value to round-off over three
    steps."""
    sys.path.insert(0, ROOT)
    import bench
    from hipft_amd import tune
    wl = bench.make_workload(64)
    g = bench.setup_model(wl)
    cfg = tune.autotune(g)  # cached on the GPU this suite runs on
    print(f"tuned launch shapes: {cfg}", flush=True)
    o = Oracle(wl["t"], wl["p"], 1)
    m = o.mesh
    Wt = np.zeros((o.npm, o.ntm))
    Wt[1:o.npm - 1, 1:o.ntm - 1] = (m["st"][1:-1] * m["dt"][1:-1])[None, :] * \
        m["dp"][1:o.npm - 1][:, None]
    npole = math.pi * m["sth"][1] * m["dth"][1] / 2.0
    spole = math.pi * m["sth"][o.nt - 2] * m["dth"][o.nt - 2] / 2.0

    def signed(f):
        s = (Wt[None] * f).sum(axis=(1, 2)) + npole * f[:, 1, 0] + spole * f[:, 1, -1]
        mag = np.abs(Wt[None] * f).sum(axis=(1, 2)) + np.abs(npole * f[:, 1, 0]) + \
            np.abs(spole * f[:, 1, -1])
        return s, mag

    s0, mag0 = signed(g.get_state("tp"))
    for _ in range(3):
        g.flux_transport_step(1.0, strang=True, subcycles=0)
    s3, mag3 = signed(g.get_state("tp"))
    drift = np.abs(s3 - s0) / mag0
    print(f"signed-flux drift over 3 steps: max {drift.max():.2e} (relative to the unsigned flux)",
          flush=True)
    assert drift.max() <= 1e-12
    assert np.all(mag3 < mag0)  # diffusion cancels opposite polarities
    g.close()


def _window_model(ntm, npm1, nr, seed, clustered=False, nbip=200):
    from hipft_amd.adjoint import FlowModel
    t = W.clustered_theta(ntm, 2.0) if clustered else W.uniform_theta(ntm)
    p = W.uniform_phi(npm1)
    vt_unit = np.stack([W.meridional_profile(t, 1.0) for _ in range(nr)])
    dr_unit = np.array([W.dr_coefficients(1.0) for _ in range(nr)])
    # default launch shapes: tuning a new grid and batch here would take minutes
    model = FlowModel(t, p, nr, vt_unit=vt_unit, dr_unit=dr_unit, autotune=False)
    rng = np.random.default_rng(seed)
    maps = np.stack([W.bipole_map(t, p, nbip, seed + 1 + r) for r in range(nr)])
    model.ft.set_state(maps, layout="pt")
    model.bind_stream()
    eta = torch.tensor(np.exp(rng.uniform(math.log(100.0), math.log(1000.0), nr)),
                       dtype=torch.float64)
    v0 = torch.tensor(rng.uniform(10.0, 20.0, nr), dtype=torch.float64)
    om = torch.tensor(rng.uniform(0.9, 1.1, nr), dtype=torch.float64)
    return model, eta, v0, om, rng


def test_config4_adjoint_central_differences(gpu_available):
    """The adjoint configuration as stated: 256x512, 16 realizations, 200 steps
    of 1 h, checkpoints every 20 steps with recompute, target = a forward run
    with eta x 1.2 plus N(0, 2 G); the gradient with respect to the initial
    maps on 32 random directions and with respect to eta, v0 and the rotation
    scale of two realizations agrees with central differences of the forward
    HIP run (h = 1e-6 relative on the map, 1e-5 on the scalars) to 1e-6."""
    from hipft_amd.adjoint import run
    nsteps, every = 200, 20
    model, eta, v0, om, rng = _window_model(256, 512, 16, 77)
    f0 = model.get()
    with torch.no_grad():
        tgt = run(f0, eta * 1.2, v0, om, model, nsteps, checkpoint_every=nsteps)
        tgt = tgt + torch.as_tensor(rng.normal(0.0, 2.0, tuple(tgt.shape)), device="cuda")
    x = f0.clone().requires_grad_(True)
    e, v, o = (a.clone().requires_grad_(True) for a in (eta, v0, om))
    y = run(x, e, v, o, model, nsteps, checkpoint_every=every)
    loss = 0.5 * ((y - tgt) ** 2).sum(dim=(1, 2))  # per realization
    gx, ge, gv, go = torch.autograd.grad(loss.sum(), (x, e, v, o))

    def L(xx, ee, vv, oo, r=None):
        with torch.no_grad():
            d = run(xx, ee, vv, oo, model, nsteps, checkpoint_every=nsteps) - tgt
            j = 0.5 * (d * d).sum(dim=(1, 2))
            return float(j.sum() if r is None else j[r])

    scale = float(f0.abs().max())
    h = 1e-6
    worst = 0.0
    for _ in range(32):
        d = torch.as_tensor(rng.standard_normal(tuple(f0.shape)), device="cuda") * scale
        fd = (L(f0 + h * d, eta, v0, om) - L(f0 - h * d, eta, v0, om)) / (2 * h)
        ad = float((gx * d).sum())
        worst = max(worst, abs(fd - ad) / abs(ad))
    print(f"\n32 map directions: worst relative difference {worst:.2e}", flush=True)
    assert worst <= 1e-6
    h = 1e-5
    for r in (0, 9):
        for k, (base, g) in enumerate(((eta, ge), (v0, gv), (om, go))):
            dv = torch.zeros(16, dtype=torch.float64)
            dv[r] = h * float(base[r])
            a_p, a_m = [eta, v0, om], [eta, v0, om]
            a_p[k], a_m[k] = base + dv, base - dv
            fd = (L(f0, *a_p, r=r) - L(f0, *a_m, r=r)) / (2 * float(dv[r]))
            assert abs(fd - float(g[r])) <= 1e-6 * abs(float(g[r])), (r, k, fd, float(g[r]))
    model.ft.close()


def test_config5_shape_window_gradient(gpu_available):
    from hipft_amd.adjoint import run
    free0, total = torch.cuda.mem_get_info()
    # 20 bipoles a map: the gradient check does not need 200
    model, eta, v0, om, rng = _window_model(512, 1024, 128, 350, clustered=True, nbip=20)
    f0 = model.get()
    with torch.no_grad():
        tgt = run(f0, torch.full_like(eta, 350.0), v0, om, model, 2, checkpoint_every=2)
    e = eta.clone().requires_grad_(True)
    y = run(f0, e, v0, om, model, 2, checkpoint_every=1)
    loss = 0.5 * ((y - tgt) ** 2).sum()
    (ge,) = torch.autograd.grad(loss, (e,))
    torch.cuda.synchronize()
    used_gb = (free0 - torch.cuda.mem_get_info()[0]) / 2 ** 30
    print(f"\n512x1024 x 128 realizations, two-step window: device memory held {used_gb:.1f} GiB")
    assert used_gb < 0.5 * total / 2 ** 30

    def L(ee):
        with torch.no_grad():
            return 0.5 * float(((run(f0, ee, v0, om, model, 2, checkpoint_every=2)
                                 - tgt) ** 2).sum())

    h = 1e-5
    for r in (0, 127):
        dv = torch.zeros(128, dtype=torch.float64)
        dv[r] = h * float(eta[r])
        fd = (L(eta + dv) - L(eta - dv)) / (2 * float(dv[r]))
        assert abs(fd - float(ge[r])) <= 1e-6 * abs(float(ge[r])), (r, fd, float(ge[r]))
    model.ft.close()


def test_config5_window_fit_and_gradients(gpu_available):
    """The gradient-assimilation configuration at its shape (512x1024, 128
    realizations) on a three-observation window: observations every 6 h of a
    truth run (eta = 350 km^2/s) with 5 G noise, mu^4-weighted within 60
    degrees of disk centre (the full 30-day window of daily observations is
    bench.py --cfg5; this keeps the test to about a minute); three Adam
    iterations on eta, v0 and the initial map lower the summed misfit, and the
    adjoint gradients of one realization's misfit with respect to its eta, v0
    and initial map match central differences of the forward HIP run
    (h = 1e-6 relative on eta, 1e-3 on v0, 1e-4 on the map, agreement 1e-6)."""
    from hipft_amd.variational import fit, synthetic_window, window_misfit
    free0, total = torch.cuda.mem_get_info()
    win = synthetic_window(512, 1024, 128, 3, nbip=20, steps_per_day=6)
    model, obs, wts, spo = win["model"], win["obs"], win["weights"], win["steps_per_obs"]
    res = fit(model, win["f0"], win["eta0"], win["v0_0"], win["omega0"], obs, wts, spo,
              iters=3, lr_log=0.05, lr_map=1.0)
    print(f"\nwindow misfit over 3 Adam iterations: {res['loss']}", flush=True)
    assert res["loss"][2] < res["loss"][1] < res["loss"][0]
    torch.cuda.synchronize()
    print(f"device memory held: {(free0 - torch.cuda.mem_get_info()[0]) / 2 ** 30:.1f} GiB",
          flush=True)
    # gradients at the fitted point (realizations are independent, so the
    # gradient of the summed misfit restricted to realization r is that of J_r)
    r = 5
    f0 = res["f0"]
    eta, v0, om = res["eta"].clone(), res["v0"].clone(), res["omega"].clone()
    x = f0.clone().requires_grad_(True)
    e, v = eta.clone().requires_grad_(True), v0.clone().requires_grad_(True)
    J = window_misfit(model, x, e, v, om, obs, wts, spo)
    gx, ge, gv = torch.autograd.grad(J.sum(), (x, e, v))

    from hipft_amd.adjoint import run

    def states(xx, ee, vv):
        with torch.no_grad():
            x, out = xx, []
            for _ in obs:
                x = run(x, ee, vv, om, model, spo)
                out.append(x[r].clone())
            return out

    def dJ(plus, minus):
        # J_r(+) - J_r(-) from the two trajectories without subtracting the
        # two large misfits (their summation rounding would swamp a central
        # difference): 1/2 sum w (x+ - x-)(x+ + x- - 2 o)
        tot = 0.0
        for sp, sm, o, w in zip(plus, minus, obs, wts):
            tot += float((0.5 * w * (sp - sm) * (sp + sm - 2.0 * o)).sum())
        return tot

    # step sizes: the misfit is dominated by the 5 G observation noise, so the
    # rounding of a forward run, dotted with that noise, puts an error of
    # ~1e-9 G^2 into a misfit difference whatever h is; it falls as 1/h in the
    # central difference.  The v0 derivative is small against that (measured
    # at 512x1024: h = 1e-6 relative leaves 1e-5 of it, h = 1e-3 1e-9), so
    # v0 takes h = 1e-3, where the truncation error (h^2) is still below 1e-8;
    # eta takes 1e-6, since a relative change of 1e-3 in eta can move the RKG2
    # stage count (a discrete choice the gradient holds fixed)
    for which, base, g, h in (("eta", eta, ge, 1e-6), ("v0", v0, gv, 1e-3)):
        dv = torch.zeros(128, dtype=torch.float64)
        dv[r] = h * float(base[r])
        args_p = [f0, eta, v0]
        args_m = [f0, eta, v0]
        k = 1 if which == "eta" else 2
        args_p[k] = base + dv
        args_m[k] = base - dv
        fd = dJ(states(*args_p), states(*args_m)) / (2 * float(dv[r]))
        assert abs(fd - float(g[r])) <= 1e-6 * abs(float(g[r])), (which, fd, float(g[r]))
    rng = np.random.default_rng(9)
    d = torch.zeros_like(f0)
    d[r] = torch.as_tensor(rng.standard_normal(tuple(f0.shape[1:])), device="cuda") * 100.0
    h = 1e-4
    fd = dJ(states(f0 + h * d, eta, v0), states(f0 - h * d, eta, v0)) / (2 * h)
    ad = float((gx * d).sum())
    assert abs(fd - ad) <= 1e-6 * abs(ad), ("map", fd, ad)
    model.ft.close()
