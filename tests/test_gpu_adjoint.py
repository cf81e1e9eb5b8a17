"""GPU discrete adjoint (hft_step_adjoint, hipft_amd.adjoint) against
(1) PyTorch autograd of the FP64 CPU transcription of the same step
    (oracle/torch_step.py), and
(2) central finite differences of the forward HIP step.
"""
import numpy as np
import pytest
import torch

from hipft_amd import workloads as W
from hipft_amd.core import FluxTransport, DIFF_KM2_S_TO_RS2_HR
from oracle.oracle import Oracle
from oracle.torch_step import TorchStep

pytestmark = pytest.mark.gpu


def build(ntm, npm1, nr, eta, flow=True, diffusion=True, seed=5, frozen=False):
    t, p = W.clustered_theta(ntm, 2.0), W.uniform_phi(npm1)
    maps = np.stack([W.bipole_map(t, p, 8, seed + r) for r in range(nr)])
    vt = np.stack([W.meridional_profile(t, 12.0 + 4.0 * r) for r in range(nr)])
    p0, p2, p4 = (np.array([c * (1.0 + 0.05 * r) for r in range(nr)]) for c in W.dr_coefficients())
    o = Oracle(t, p, nr)
    o.ingest_pt(maps)
    g = FluxTransport(t, p, nr, mode="independent")
    g.set_stream(torch.cuda.current_stream().cuda_stream)
    g.set_state(maps, layout="pt")
    if flow:
        o.flow_reset()
        o.flow_set(np.ascontiguousarray(np.broadcast_to(vt[:, None, :], (nr, o.npm, o.nt))),
                   np.zeros((nr, o.np_, o.ntm)))
        o.flow_add_dr(p0, p2, p4)
        o.set_lf_alpha()
        g.set_flow_profiles(vt, None)
        g.set_flow_analytic(p0=p0, p2=p2, p4=p4)
    if diffusion:
        o.load_diffusion(list(eta))
        g.set_diffusion(eta)
    ts = TorchStep(o, D=np.asarray(eta) * DIFF_KM2_S_TO_RS2_HR if diffusion else None,
                   flow=flow, diffusion=diffusion, frozen_weno=frozen)
    return o, g, ts


def dev_state(g):
    x = torch.empty((g.nr, g.ntm, g.npm - 2), dtype=torch.float64, device="cuda")
    g.get_state_device(x.data_ptr(), g.npm - 2)
    return x


def rel(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


def gpu_step_and_adjoint(g, x0, tgt, dt, strang, subcycles, frozen=False):
    nu = g.npm - 2
    g.set_state_device(x0.data_ptr(), nu)
    g.flux_transport_step(dt, strang=strang, subcycles=subcycles)
    y = dev_state(g)
    lam = (y - tgt).contiguous()
    g.set_state_device(x0.data_ptr(), nu)
    grads = g.step_adjoint(dt, lam.data_ptr(), nu, strang=strang, subcycles=subcycles,
                           frozen_weno=frozen)
    torch.cuda.synchronize()
    return y, lam, grads


def torch_grads(ts, x0, tgt, dt, strang, subcycles, s_r):
    nr = ts.nr
    x = x0.detach().cpu().clone().requires_grad_(True)
    logD = torch.zeros(nr, dtype=torch.float64, requires_grad=True)
    st = torch.ones(nr, dtype=torch.float64, requires_grad=True)
    sp = torch.ones(nr, dtype=torch.float64, requires_grad=True)
    y = ts.step(x, logD, st, sp, dt, strang=strang, subcycles=subcycles, s_r=s_r)
    J = 0.5 * ((y - tgt.cpu()) ** 2).sum()
    J.backward()
    return y.detach(), x.grad, logD.grad, st.grad, sp.grad


@pytest.mark.parametrize("what", ["diffusion", "advection", "step"])
def test_step_adjoint_matches_autograd(gpu_available, what):
    flow = what in ("advection", "step")
    diffusion = what in ("diffusion", "step")
    eta = [150.0, 420.0, 900.0]
    o, g, ts = build(33, 65, 3, eta, flow=flow, diffusion=diffusion)
    x0 = dev_state(g)
    rng = np.random.default_rng(11)
    tgt = torch.as_tensor(rng.standard_normal(tuple(x0.shape)), device="cuda")
    s_r = ts.stage_counts(1.0) if diffusion else None
    y, lam, grads = gpu_step_and_adjoint(g, x0, tgt, 1.0, True, 1)
    yt, gx, gD, gst, gsp = torch_grads(ts, x0, tgt, 1.0, True, 1, s_r)
    assert rel(y, yt) < 1e-12, "forward step"
    assert rel(lam, gx) < 1e-10, f"dJ/dx ({what})"
    if diffusion:
        assert rel(grads[:, 0], gD) < 1e-10, "dJ/dlog eta"
    if flow:
        assert rel(grads[:, 1], gst) < 1e-10, "dJ/dlog vt scale"
        assert rel(grads[:, 2], gsp) < 1e-10, "dJ/dlog vp scale"


@pytest.mark.parametrize("shape", [(33, 1025), (200, 129)])
@pytest.mark.parametrize("what", ["advection", "step"])
def test_step_adjoint_matches_autograd_large_grids(gpu_available, shape, what):
    """The same comparison where the kernels take their other shapes: more
    than 512 phi columns (the WENO3 transpose in 1024-thread blocks, the
    fused RKG2 launches over the full width) and 200 theta rows (13 bands of
    the WENO3 transpose, several fused bands)."""
    diffusion = what == "step"
    eta = [150.0, 420.0, 900.0]
    o, g, ts = build(*shape, 3, eta, flow=True, diffusion=diffusion)
    x0 = dev_state(g)
    rng = np.random.default_rng(12)
    tgt = torch.as_tensor(rng.standard_normal(tuple(x0.shape)), device="cuda")
    s_r = ts.stage_counts(1.0) if diffusion else None
    y, lam, grads = gpu_step_and_adjoint(g, x0, tgt, 1.0, True, 1)
    yt, gx, gD, gst, gsp = torch_grads(ts, x0, tgt, 1.0, True, 1, s_r)
    assert rel(y, yt) < 1e-12, "forward step"
    assert rel(lam, gx) < 1e-10, f"dJ/dx ({what})"
    if diffusion:
        assert rel(grads[:, 0], gD) < 1e-10, "dJ/dlog eta"
    assert rel(grads[:, 1], gst) < 1e-10, "dJ/dlog vt scale"
    assert rel(grads[:, 2], gsp) < 1e-10, "dJ/dlog vp scale"


def test_frozen_weights_adjoint_matches_frozen_autograd(gpu_available):
    """Linearised WENO3 with frozen smoothness weights, checked separately
    against autograd of the transcription with detached weights."""
    o, g, ts = build(33, 65, 2, [300.0, 300.0], diffusion=False, frozen=True)
    x0 = dev_state(g)
    tgt = torch.zeros_like(x0)
    _, lam, grads = gpu_step_and_adjoint(g, x0, tgt, 1.0, False, 1, frozen=True)
    _, gx, _, gst, gsp = torch_grads(ts, x0, tgt, 1.0, False, 1, None)
    assert rel(lam, gx) < 1e-10
    assert rel(grads[:, 1], gst) < 1e-10
    assert rel(grads[:, 2], gsp) < 1e-10


@pytest.mark.parametrize("fused_min,T", [(1, 4), (8, 4), (1, 5)])
def test_segmented_rkg2_adjoint_matches_stored(gpu_available, fused_min, T):
    """Segment recomputation of RKG2 stages (short segments, one-stage
    launches) gives the same adjoint as storing every stage during the
    forward recomputation (fused multi-stage launches of T stages, each
    stage also written out, when fused_min = 1)."""
    o, g, ts = build(41, 81, 3, [120.0, 500.0, 1000.0], flow=False)
    g.set_tuning(fused=1, fused_kernel=3, fused_stages=T, band_rows=16,
                 fused_min_realizations=fused_min)
    x0 = dev_state(g)
    tgt = torch.zeros_like(x0)
    out = []
    for seg in (0, 3, 7):
        g.set_tuning(adjoint_segment=seg)
        _, lam, grads = gpu_step_and_adjoint(g, x0, tgt, 2.0, True, 2)
        out.append((lam.cpu(), grads))
    for lam, grads in out[1:]:
        assert rel(lam, out[0][0]) < 1e-13
        assert np.allclose(grads, out[0][1], rtol=1e-12, atol=0)


@pytest.mark.parametrize("shape", [(65, 129, 16, 12.0), (67, 700, 8, 1.0), (130, 257, 32, 4.0),
                                   (97, 193, 4, 4.0), (41, 81, 64, 12.0)])
def test_fused_rkg2_adjoint_matches_one_stage_launches(gpu_available, shape):
    """The multi-stage RKG2 adjoint launch over stored stages (T stages per
    launch, realizations with different stage counts entering together and
    finishing inside launches) gives the adjoint of the one-stage launches:
    bands of 4..32 rows, one band over the whole grid, a grid wider than 512
    phi columns (1024-thread blocks).  Stage counts (oracle): 7/13/19,
    11/21/31, 15/31/43, 9/17/25, 3/5/7."""
    ntm, npm1, rows, dt = shape
    print(f"fused RKG2 adjoint: {ntm}x{npm1}, {rows}-row bands, dt {dt}", flush=True)
    o, g, ts = build(ntm, npm1, 3, [120.0, 500.0, 1000.0], flow=False)
    x0 = dev_state(g)
    rng = np.random.default_rng(3)
    tgt = torch.as_tensor(rng.standard_normal(tuple(x0.shape)), device="cuda")
    out = []
    for fused in (0, 1):
        g.set_tuning(adjoint_fused=fused, adjoint_fused_rows=rows)
        _, lam, grads = gpu_step_and_adjoint(g, x0, tgt, dt, False, 1)
        out.append((lam.cpu(), grads))
    s_r = g.diffusion_stats()[0]
    assert len(set(int(s) for s in s_r)) == 3, f"three different stage counts {s_r}"
    assert rel(out[1][0], out[0][0]) < 1e-13
    assert np.allclose(out[1][1][:, 0], out[0][1][:, 0], rtol=1e-12, atol=0)
    g.close()


@pytest.mark.parametrize("mem_mb", [32768, 1])
def test_multi_step_adjoint_matches_step_by_step(gpu_available, mem_mb):
    """hft_steps_adjoint over 3 steps (one storing forward pass, or step
    inputs with recomputation when the memory budget is tiny) equals three
    hft_step_adjoint calls in reverse, and leaves the final state."""
    o, g, ts = build(33, 65, 2, [250.0, 800.0])
    nu = g.npm - 2
    x0 = dev_state(g)
    xs = [x0]
    for _ in range(3):
        g.flux_transport_step(1.0, strang=True, subcycles=2)
        xs.append(dev_state(g))
    rng = np.random.default_rng(3)
    lam0 = torch.as_tensor(rng.standard_normal(tuple(x0.shape)), device="cuda")
    lam = lam0.clone()
    ref = np.zeros((2, 3))
    for n in (2, 1, 0):
        g.set_state_device(xs[n].data_ptr(), nu)
        ref += g.step_adjoint(1.0, lam.data_ptr(), nu, strang=True, subcycles=2)
    g.set_tuning(adjoint_mem_mb=mem_mb)
    g.set_state_device(x0.data_ptr(), nu)
    lam2 = lam0.clone()
    got = g.steps_adjoint(3, 1.0, lam2.data_ptr(), nu, strang=True, subcycles=2)
    torch.cuda.synchronize()
    assert rel(lam2, lam) < 1e-13
    assert np.allclose(got, ref, rtol=1e-12, atol=0)
    assert rel(dev_state(g), xs[3]) == 0.0


def test_gradient_vs_central_differences(gpu_available):
    """32 random map directions (h = 1e-6 relative) and the three parameters,
    against the forward HIP step (two steps, fixed sub-cycling).  The target
    is the state reached with perturbed parameters, so the misfit is not
    dominated by a parameter-independent constant (whose rounding would swamp
    a central difference at h = 1e-6)."""
    eta = np.array([200.0, 700.0])
    o, g, ts = build(33, 65, 2, eta)
    nu = g.npm - 2
    x0 = dev_state(g)
    rng = np.random.default_rng(7)
    vt = np.stack([W.meridional_profile(o.mesh["t"], 12.0 + 4.0 * r) for r in range(2)])
    dr = np.array([W.dr_coefficients(1.0 + 0.05 * r) for r in range(2)])
    one = np.ones(2)

    def set_par(e, sv, so):
        g.set_diffusion(e)
        g.set_flow_profiles(vt * sv[:, None], None)
        g.set_flow_analytic(p0=dr[:, 0] * so, p2=dr[:, 1] * so, p4=dr[:, 2] * so)

    def fwd(x):
        g.set_state_device(x.data_ptr(), nu)
        for _ in range(2):
            g.flux_transport_step(1.0, strang=True, subcycles=1)
        return dev_state(g)

    set_par(eta * 1.2, one * 1.1, one * 0.9)
    tgt = fwd(x0)
    set_par(eta, one, one)

    def J(x):
        return 0.5 * float(((fwd(x) - tgt) ** 2).sum())

    # adjoint over two steps
    g.set_state_device(x0.data_ptr(), nu)
    g.flux_transport_step(1.0, strang=True, subcycles=1)
    x1 = dev_state(g)
    g.flux_transport_step(1.0, strang=True, subcycles=1)
    lam = (dev_state(g) - tgt).contiguous()
    grads = np.zeros((2, 3))
    for x in (x1, x0):
        g.set_state_device(x.data_ptr(), nu)
        grads += g.step_adjoint(1.0, lam.data_ptr(), nu, strang=True, subcycles=1)
    scale = float(x0.abs().max())
    worst = 0.0
    for _ in range(32):
        d = torch.as_tensor(rng.standard_normal(tuple(x0.shape)), device="cuda") * scale
        h = 1e-6
        fd = (J(x0 + h * d) - J(x0 - h * d)) / (2 * h)
        ad = float((lam * d).sum())
        worst = max(worst, abs(fd - ad) / abs(ad))
    assert worst <= 1e-6, worst

    def J_par(e, sv, so):
        set_par(e, sv, so)
        return J(x0)

    # parameters: h = 1e-5 relative (the velocity-scale sensitivities are
    # ~1e-5 of J, so at 1e-6 the rounding of J is ~1e-6 of the difference)
    h = 1e-5
    for r in range(2):
        e1 = np.zeros(2)
        e1[r] = 1.0
        fd = (J_par(eta * (1 + h * e1), one, one) - J_par(eta * (1 - h * e1), one, one)) / (2 * h)
        assert abs(fd - grads[r, 0]) <= 1e-6 * abs(grads[r, 0]), (r, fd, grads[r, 0])
        fd = (J_par(eta, one + h * e1, one) - J_par(eta, one - h * e1, one)) / (2 * h)
        assert abs(fd - grads[r, 1]) <= 1e-6 * abs(grads[r, 1]), (r, fd, grads[r, 1])
        fd = (J_par(eta, one, one + h * e1) - J_par(eta, one, one - h * e1)) / (2 * h)
        assert abs(fd - grads[r, 2]) <= 1e-6 * abs(grads[r, 2]), (r, fd, grads[r, 2])


def test_autograd_function_with_checkpoints(gpu_available):
    """torch.autograd.Function over several steps (checkpoint every 2 steps)
    equals the gradient with a checkpoint at every step, and its parameter
    gradients match central differences of the loss."""
    from hipft_amd.adjoint import FlowModel, run
    ntm, npm1, nr = 33, 65, 2
    t, p = W.clustered_theta(ntm, 2.0), W.uniform_phi(npm1)
    vt_unit = np.stack([W.meridional_profile(t, 1.0) for _ in range(nr)])
    dr_unit = np.array([W.dr_coefficients(1.0) for _ in range(nr)])
    model = FlowModel(t, p, nr, vt_unit=vt_unit, dr_unit=dr_unit)
    maps = np.stack([W.bipole_map(t, p, 8, 40 + r) for r in range(nr)])
    model.ft.set_state(maps, layout="pt")
    model.bind_stream()
    f0 = model.get()
    eta = torch.tensor([250.0, 600.0], dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor([12.0, 18.0], dtype=torch.float64, requires_grad=True)
    om = torch.tensor([0.95, 1.05], dtype=torch.float64, requires_grad=True)
    with torch.no_grad():  # target: the run with perturbed parameters
        tgt = run(f0, eta * 1.2, v0 * 1.1, om * 0.9, model, 5, dt=1.0)
    res = []
    for every in (1, 2):
        x = f0.clone().requires_grad_(True)
        y = run(x, eta, v0, om, model, 5, dt=1.0, checkpoint_every=every)
        loss = 0.5 * ((y - tgt) ** 2).sum()
        gx, ge, gv, go = torch.autograd.grad(loss, (x, eta, v0, om))
        res.append((float(loss), gx, ge, gv, go))
    assert rel(res[1][1], res[0][1]) < 1e-13
    for k in (2, 3, 4):
        assert rel(res[1][k], res[0][k]) < 1e-12

    def L(e, v, o_):
        with torch.no_grad():
            y = run(f0, e, v, o_, model, 5, dt=1.0)
            return 0.5 * float(((y - tgt) ** 2).sum())

    h = 1e-6
    with torch.no_grad():
        for k, (base, args) in enumerate(((eta, 0), (v0, 1), (om, 2))):
            dv = torch.zeros(2, dtype=torch.float64)
            dv[0] = h * float(base[0])
            a_p = [eta.detach(), v0.detach(), om.detach()]
            a_m = list(a_p)
            a_p[args] = base.detach() + dv
            a_m[args] = base.detach() - dv
            fd = (L(*a_p) - L(*a_m)) / (2 * float(dv[0]))
            ad = float(res[0][2 + k][0])
            assert abs(fd - ad) <= 1e-6 * abs(ad), (k, fd, ad)


def test_gradient_assimilation_recovers_diffusivity(gpu_available):
    """Batched gradient assimilation (hipft_amd.variational.fit): each of four
    realizations, started from a wrong diffusivity, fits its own window of
    observed maps from a truth run (eta = 350 km^2/s) with Adam on log eta;
    the window misfit falls and every eta moves to the truth."""
    from hipft_amd.adjoint import FlowModel, run
    from hipft_amd.variational import fit
    ntm, npm1, nr = 33, 65, 4
    t, p = W.uniform_theta(ntm), W.uniform_phi(npm1)
    vt_unit = np.stack([W.meridional_profile(t, 1.0) for _ in range(nr)])
    dr_unit = np.array([W.dr_coefficients(1.0) for _ in range(nr)])
    model = FlowModel(t, p, nr, vt_unit=vt_unit, dr_unit=dr_unit)
    maps = np.stack([W.bipole_map(t, p, 10, 90 + r) for r in range(nr)])
    model.ft.set_state(maps, layout="pt")
    model.bind_stream()
    f0 = model.get()
    truth = torch.full((nr,), 350.0, dtype=torch.float64)
    v0 = torch.full((nr,), 15.0, dtype=torch.float64)
    om = torch.ones(nr, dtype=torch.float64)
    obs = []
    with torch.no_grad():
        x = f0
        for _ in range(4):
            x = run(x, truth, v0, om, model, 6)
            obs.append(x.clone())
    eta0 = torch.tensor([150.0, 250.0, 500.0, 700.0], dtype=torch.float64)
    res = fit(model, f0, eta0, v0, om, obs, None, 6, iters=40, fit_map=False, fit_v0=False,
              lr_log=0.1)
    assert res["loss"][-1] < 1e-2 * res["loss"][0], res["loss"][::10]
    err0 = torch.log(eta0 / 350.0).abs()
    err1 = torch.log(res["eta"] / 350.0).abs()
    assert bool((err1 < 0.1 * err0).all()), (res["eta"], err1)
