"""GPU: flux-transport steps planned on the device and replayed as one hipGraph
(tuning key step_graph) against the host-controlled step and the CPU oracle.

The planned step evaluates the diffusion controller on the device: the PTL
step rule, the RKG2 stage counts and coefficient tables and the sub-cycle
bookkeeping (diffusion_step, get_dtime_diffusion_ptl and load_sts_rkg2,
src/hipft.f90:4589-4690, 5077-5156, 5399-5461).  The kernels that update the
maps are the host path's, so every state must equal the host-controlled
step's bit for bit, and the oracle's.
"""
import numpy as np
import pytest

from hipft_amd import workloads as W
from hipft_amd.core import FluxTransport
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def grid(ntm, npm1):
    return W.clustered_theta(ntm, 2.0), W.uniform_phi(npm1)


def maps_for(t, p, nr, seed, noise):
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nr):
        m = W.bipole_map(t, p, 20, seed + r)
        if noise:
            # rough maps: the PTL limit cuts a step into several cycles
            m = m + noise * rng.standard_normal(m.shape)
        out.append(m)
    return np.stack(out)


def model(t, p, maps, eta, mode, step_graph, **tuning):
    nr = len(eta)
    g = FluxTransport(t, p, nr, mode=mode)
    g.set_state(maps, layout="pt")
    g.set_diffusion(eta)
    vt = np.stack([W.meridional_profile(t, 10.0 + 2.0 * r) for r in range(nr)])
    g.set_flow_profiles(vt, None)
    dr = np.array([W.dr_coefficients(0.9 + 0.05 * r) for r in range(nr)])
    g.set_flow_analytic(p0=dr[:, 0], p2=dr[:, 1], p4=dr[:, 2])
    g.set_tuning(step_graph=step_graph, **tuning)
    return g


def same(a, b, what):
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} entries differ, first {bad[:3].tolist()}, " \
                          f"max |diff| {np.abs(a - b).max():.3e}"


def lockstep(h, g, steps, dt, subcycles=0, strang=True):
    """Advance the host-controlled model h and the planned model g together,
    comparing states and diffusion statistics after every step."""
    cycles = []
    for i in range(steps):
        h.flux_transport_step(dt, strang=strang, subcycles=subcycles)
        g.flux_transport_step(dt, strang=strang, subcycles=subcycles)
        same(g.get_state("tp"), h.get_state("tp"), f"state after step {i + 1}")
        sh, ch = h.diffusion_stats()
        sg, cg = g.diffusion_stats()
        same(sg, sh, f"stage counts after step {i + 1}")
        same(cg, ch, f"cycle counts after step {i + 1}")
        cycles.append(int(ch.max()))
    return cycles


@pytest.mark.parametrize("mode,nr", [("independent", 6), ("coupled", 4)])
def test_planned_step_matches_host_step(gpu_available, mode, nr):
    """PTL sub-cycling, Strang steps, per-realization stage counts (independent)
    or one controller for the batch (coupled): the first step is captured,
    later ones replay the captured sequence."""
    t, p = grid(91, 181)
    eta = np.linspace(120.0, 950.0, nr)
    maps = maps_for(t, p, nr, 300, noise=0.0)
    h = model(t, p, maps, eta, mode, 0)
    g = model(t, p, maps, eta, mode, 1)
    lockstep(h, g, 6, 1.0)
    assert g.get_tuning("step_graph_builds") == 1
    assert g.get_tuning("step_graph_hits") == 5
    assert h.get_tuning("step_graph_builds") == 0
    h.close()
    g.close()


def test_planned_step_fixed_subcycles(gpu_available):
    """diffusion_subcycles > 0: every round has the same stage count, so the
    sequence is known before the first step and no round overflows."""
    t, p = grid(91, 181)
    eta = np.array([150.0, 420.0, 880.0])
    maps = maps_for(t, p, 3, 310, noise=0.0)
    h = model(t, p, maps, eta, "independent", 0)
    g = model(t, p, maps, eta, "independent", 1)
    lockstep(h, g, 4, 2.0, subcycles=3)
    assert g.get_tuning("step_graph_builds") == 1
    assert g.get_tuning("step_graph_hits") == 3
    assert g.get_tuning("step_graph_overflows") == 0
    h.close()
    g.close()


@pytest.mark.parametrize("rounds", [1, 0])
def test_planned_step_overflow_resumes_on_host(gpu_available, rounds):
    """Rough maps make the PTL rule cut each step into several cycles.  With
    the captured sequence held to one round (rounds=1, the default) a step of
    several cycles overflows after its first round and the host controller
    finishes it from the device's controller state; with learned rounds
    (rounds=0) the sequence grows with the rounds seen.  Both equal the host
    path bit for bit."""
    t, p = grid(91, 181)
    eta = np.array([120.0, 480.0, 950.0])
    maps = maps_for(t, p, 3, 320, noise=5.0)
    h = model(t, p, maps, eta, "independent", 0)
    g = model(t, p, maps, eta, "independent", 1, step_graph_rounds=rounds)
    cycles = lockstep(h, g, 6, 1.0)
    if rounds == 1:
        # one captured round, sized for every stage count a step can need:
        # every step is planned (one capture), only extra rounds overflow
        multi = sum(1 for c in cycles if c >= 2)
        assert multi >= 2, cycles
        assert g.get_tuning("step_graph_builds") == 1
        assert g.get_tuning("step_graph_hits") == 5
        assert g.get_tuning("step_graph_overflows") == multi, cycles
    else:
        # the first step learns the rounds; every later one is planned
        assert max(cycles[1:]) >= 2, cycles
        assert g.get_tuning("step_graph_hits") + g.get_tuning("step_graph_builds") == 5
    h.close()
    g.close()


def test_planned_step_matches_oracle(gpu_available):
    """Planned steps of a batch of three against single-realization oracle runs
    (independent controllers), bit for bit, over three Strang steps (capture,
    then replays)."""
    t, p = grid(41, 81)
    nr = 3
    eta = np.array([250.0, 700.0, 130.0])
    maps = maps_for(t, p, nr, 330, noise=0.0)
    g = model(t, p, maps, eta, "independent", 1)
    vt = np.stack([W.meridional_profile(t, 10.0 + 2.0 * r) for r in range(nr)])
    dr = np.array([W.dr_coefficients(0.9 + 0.05 * r) for r in range(nr)])
    oracles = []
    for r in range(nr):
        o = Oracle(t, p, 1)
        o.ingest_pt(maps[r])
        vt2 = np.ascontiguousarray(np.broadcast_to(vt[r][None, None, :], (1, o.npm, o.nt)))
        o.flow_reset()
        o.flow_set(vt2, np.zeros((1, o.np_, o.ntm)))
        o.flow_add_dr([dr[r, 0]], [dr[r, 1]], [dr[r, 2]])
        o.set_lf_alpha()
        o.load_diffusion([eta[r]])
        oracles.append(o)
    for i in range(3):
        g.flux_transport_step(1.0, strang=True, subcycles=0)
        fg = g.get_state("tp")
        for r, o in enumerate(oracles):
            o.advection_step(0.5)
            o.diffusion_step(1.0, 0)
            o.advection_step(0.5)
            same(fg[r], o.get_f()[0], f"step {i + 1}, realization {r}")
    assert g.get_tuning("step_graph_hits") >= 1
    g.close()


def test_planned_step_benchmark_batch(gpu_available):
    """The 64-realization 512x1024 benchmark batch with its tuned launch
    shapes: planned steps equal host-controlled steps bit for bit (capture,
    three replays)."""
    import bench
    from hipft_amd import tune
    wl = bench.make_workload(64)
    h = bench.setup_model(wl)
    g = bench.setup_model(wl)
    tune.autotune(h, write_cache=False, borrow=True)
    tune.autotune(g, write_cache=False, borrow=True)
    h.set_tuning(step_graph=0)
    g.set_tuning(step_graph=1)
    lockstep(h, g, 4, 1.0)
    assert g.get_tuning("step_graph_builds") == 1
    assert g.get_tuning("step_graph_hits") == 3
    h.close()
    g.close()


def test_planned_step_attenuated_flows_updated_each_step(gpu_available):
    """Flows attenuated by |Br| (2-D flow arrays) recomputed from the state
    before every step, as the driver does: the captured step reads the flow
    arrays in place, so it is captured once and replayed, bit for bit with
    host-controlled steps (tanh of the attenuation is the same device code on
    both paths)."""
    t, p = grid(91, 181)
    eta = np.array([200.0, 600.0])
    maps = maps_for(t, p, 2, 340, noise=0.0)
    h = model(t, p, maps, eta, "independent", 0)
    g = model(t, p, maps, eta, "independent", 1)
    for m in (h, g):
        m.set_flow_attenuation([300.0, 150.0])
    for i in range(4):
        for m in (h, g):
            m.update_flow()
        h.flux_transport_step(1.0, strang=True, subcycles=0)
        g.flux_transport_step(1.0, strang=True, subcycles=0)
        same(g.get_state("tp"), h.get_state("tp"), f"state after step {i + 1}")
    assert g.get_tuning("step_graph_builds") == 1
    assert g.get_tuning("step_graph_hits") == 3
    h.close()
    g.close()


def test_planned_step_declines_cfl_substeps(gpu_available):
    """A step whose advection takes per-realization CFL sub-steps (here a
    24 h step) is not planned: it runs host-controlled, with the same
    results."""
    t, p = grid(91, 181)
    eta = np.array([200.0, 600.0, 900.0])
    maps = maps_for(t, p, 3, 350, noise=0.0)
    h = model(t, p, maps, eta, "independent", 0)
    g = model(t, p, maps, eta, "independent", 1)
    nsub = np.ceil(12.0 / g.flow_dtmax(False)[1])  # the Strang half-step of 12 h
    assert nsub.max() > 1, nsub
    lockstep(h, g, 2, 24.0)
    assert g.get_tuning("step_graph_builds") == 0
    h.close()
    g.close()


def test_planned_step_single_realization(gpu_available):
    """One realization per context (thin bands fill the device), planned and
    host-controlled, bit for bit."""
    t, p = grid(181, 361)
    eta = np.array([500.0])
    maps = maps_for(t, p, 1, 360, noise=0.0)
    h = model(t, p, maps, eta, "independent", 0, fused_target_blocks=256, band_rows=64)
    g = model(t, p, maps, eta, "independent", 1, fused_target_blocks=256, band_rows=64)
    lockstep(h, g, 3, 1.0)
    assert g.get_tuning("step_graph_builds") == 1
    assert g.get_tuning("step_graph_hits") == 2
    h.close()
    g.close()


def test_planned_step_coupled_fixed_subcycles(gpu_available):
    """Coupled controller with fixed sub-cycles (the testsuite's drop-in
    semantics with diffusion_subcycles > 0), planned and host-controlled."""
    t, p = grid(91, 181)
    eta = np.array([150.0, 420.0, 880.0, 300.0])
    maps = maps_for(t, p, 4, 370, noise=0.0)
    h = model(t, p, maps, eta, "coupled", 0)
    g = model(t, p, maps, eta, "coupled", 1)
    lockstep(h, g, 3, 3.0, subcycles=2)
    assert g.get_tuning("step_graph_builds") == 1
    assert g.get_tuning("step_graph_hits") == 2
    assert g.get_tuning("step_graph_overflows") == 0
    h.close()
    g.close()
