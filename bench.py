#!/usr/bin/env python3
"""Benchmark: realization-steps/s of the flux-transport step at 512x1024 FP64.

Workload (batched realizations): theta 512 nodes clustered x2 toward the poles
(tanh blend), phi 1024 nodes (1023 unique), 64 realizations, FP64.  Each
realization: initial map of 200 seeded bipoles (sigma log-uniform 1-4 deg,
|B| 20-500 G, Joy tilt 5 deg sin(lat)); eta log-uniform 100-1000 km^2/s;
meridional flow v0 sin(2 lat) cos(lat) with v0 uniform 10-20 m/s;
differential rotation (0.18 - 2.396 cos^2 - 1.787 cos^4) deg/day times a
scale uniform 0.9-1.1.  dt = 1 h, Strang splitting (SSPRK(4,3)+WENO3
advection half steps around RKG2 diffusion with PTL sub-cycling), every
realization with its own stage counts (independent mode), history analysis
every step.  Inputs are resident in HBM before the timed region.

A "step" is one time step of the whole batch: value = realizations x steps /
wall time of the timed steps.

Prints one JSON line.  Extra objects:
  roofline     -- dominant kernel (the multi-stage RKG2 launch): compulsory
                  bytes per launch / average launch time from HIP events on
                  the library's stream over the timed steps; traffic = HBM
                  bytes per launch from the committed rocprofv3
                  FETCH_SIZE/WRITE_SIZE summary; the stage-by-stage
                  equivalent and the whole step's measured DRAM fraction
                  beside it
  cpu_baseline -- the CPU oracle (1 core) on a bounded sample of the workload
  single       -- the same step for one realization per launch (plain path)
  adjoint      -- cost of the adjoint gradient relative to the forward run
                  (256x512, 16 realizations, 200 steps, checkpoints every 20)
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "realization-steps/s at 512x1024 FP64, %HBM roofline, adjoint/forward cost"
HBM_PEAK_GBS = 8000.0


def make_workload(nr, seed=2024, ntm=512, npm1=1024, nbip=200):
    from hipft_amd import workloads as W
    t = W.clustered_theta(ntm, 2.0)
    p = W.uniform_phi(npm1)
    rng = np.random.default_rng(seed)
    eta = np.exp(rng.uniform(math.log(100.0), math.log(1000.0), nr))
    v0 = rng.uniform(10.0, 20.0, nr)
    oms = rng.uniform(0.9, 1.1, nr)
    maps = np.stack([W.bipole_map(t, p, nbip, seed + 1 + r) for r in range(nr)])
    vt = np.stack([W.meridional_profile(t, v) for v in v0])
    dr = np.array([W.dr_coefficients(s) for s in oms])  # [nr][3]
    return dict(t=t, p=p, eta=eta, v0=v0, oms=oms, maps=maps, vt=vt, dr=dr)


def setup_model(wl, mode="independent"):
    from hipft_amd.core import FluxTransport
    nr = len(wl["eta"])
    m = FluxTransport(wl["t"], wl["p"], nr, mode=mode)
    m.set_state(wl["maps"], layout="pt")
    m.set_flow_profiles(wl["vt"], None)
    m.set_flow_analytic(p0=wl["dr"][:, 0], p2=wl["dr"][:, 1], p4=wl["dr"][:, 2])
    m.set_diffusion(wl["eta"])
    return m


def step(m, dt=1.0):
    m.flux_transport_step(dt, strang=True, subcycles=0)
    return m.analysis()


def time_steps(m, k):
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        step(m)
    m.synchronize()
    return time.perf_counter() - t0


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, budget_s=6.0, warmup=2, repeats=3):
    """The oracle (CPU restatement, OpenMP over all host threads, and one
    thread) on realization 0 of the workload: warm-up steps, then the median
    rate of `repeats` timed runs of n steps each, n sized from the warm-up so
    that a run takes about budget_s / repeats."""
    from oracle.oracle import Oracle, get_threads, set_threads
    nthreads = get_threads()

    def run(threads):
        set_threads(threads)
        o = Oracle(wl["t"], wl["p"], 1)
        o.ingest_pt(wl["maps"][0])
        nt = len(wl["t"]) + 1
        vt = np.ascontiguousarray(np.broadcast_to(wl["vt"][0][None, None, :], (1, o.npm, nt)))
        o.flow_reset()
        o.flow_set(vt, np.zeros((1, o.np_, o.ntm)))
        o.flow_add_dr(*[[c] for c in wl["dr"][0]])
        o.set_lf_alpha()
        o.load_diffusion([wl["eta"][0]])

        def one():
            o.advection_step(0.5)
            o.diffusion_step(1.0, 0)
            o.advection_step(0.5)
            o.analysis()

        t0 = time.perf_counter()
        for _ in range(warmup):
            one()
        per = (time.perf_counter() - t0) / warmup
        n = max(1, int(budget_s / repeats / per))
        rates = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            for _ in range(n):
                one()
            rates.append(n / (time.perf_counter() - t0))
        return float(np.median(rates)), n, rates

    v_all, n_all, r_all = run(nthreads)
    v_one, n_one, r_one = run(1)
    set_threads(nthreads)
    return dict(value=round(v_all, 4), unit="realization-steps/s", cores=nthreads, kind="port",
                sample=f"realization 0 (eta={wl['eta'][0]:.1f} km^2/s) of the 512x1024 workload: "
                       f"{warmup} warm-up steps, then the median of {repeats} runs of {n_all} "
                       f"steps; CPU oracle (C restatement, OpenMP over points) on {nthreads} "
                       f"threads of {os.cpu_count()} CPUs ({_cpu_model()})",
                repeats=[round(x, 4) for x in r_all],
                single_core=dict(value=round(v_one, 4), cores=1, steps_per_run=n_one,
                                 repeats=[round(x, 4) for x in r_one]))


def adjoint_bench(nr=16, ntm=256, npm1=512, nsteps=200, every=20, seed=77):
    """Gradient of a final-map misfit (256x512, 16 realizations, 200 steps of
    1 h, checkpoint every 20 steps) with respect to the initial maps and the
    per-realization eta, meridional amplitude and rotation scale.  The target
    is a forward run with eta x 1.2 plus N(0, 2 G) noise.  Returns wall times
    of the plain forward run and of forward + backward."""
    import torch
    from hipft_amd import workloads as W
    from hipft_amd.adjoint import FlowModel, run
    t, p = W.uniform_theta(ntm), W.uniform_phi(npm1)
    rng = np.random.default_rng(seed)
    eta = torch.tensor(np.exp(rng.uniform(math.log(100.0), math.log(1000.0), nr)),
                       dtype=torch.float64)
    v0 = torch.tensor(rng.uniform(10.0, 20.0, nr), dtype=torch.float64)
    om = torch.tensor(rng.uniform(0.9, 1.1, nr), dtype=torch.float64)
    vt_unit = np.stack([W.meridional_profile(t, 1.0) for _ in range(nr)])
    dr_unit = np.array([W.dr_coefficients(1.0) for _ in range(nr)])
    model = FlowModel(t, p, nr, vt_unit=vt_unit, dr_unit=dr_unit)
    maps = np.stack([W.bipole_map(t, p, 200, seed + 1 + r) for r in range(nr)])
    model.ft.set_state(maps, layout="pt")
    model.bind_stream()
    f0 = model.get()
    with torch.no_grad():
        tgt = run(f0, eta * 1.2, v0, om, model, nsteps, checkpoint_every=nsteps)
        tgt = tgt + torch.as_tensor(rng.normal(0.0, 2.0, tuple(tgt.shape)), device="cuda")
        run(f0, eta, v0, om, model, 2, checkpoint_every=nsteps)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(f0, eta, v0, om, model, nsteps, checkpoint_every=nsteps)
        torch.cuda.synchronize()
        fwd = time.perf_counter() - t0
        # the production forward: PTL sub-cycling (the differentiated step
        # takes one RKG2 cycle per step, whose length does not depend on the
        # state)
        run(f0, eta, v0, om, model, 2, checkpoint_every=nsteps, subcycles=0)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(f0, eta, v0, om, model, nsteps, checkpoint_every=nsteps, subcycles=0)
        torch.cuda.synchronize()
        fwd_ptl = time.perf_counter() - t0
    for rep in range(2):  # the first pass allocates checkpoints and adjoint buffers
        x = f0.clone().requires_grad_(True)
        e, v, o = (a.clone().requires_grad_(True) for a in (eta, v0, om))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        y = run(x, e, v, o, model, nsteps, checkpoint_every=every)
        loss = 0.5 * ((y - tgt) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    s_r = model.ft.diffusion_stats()[0]
    model.ft.close()
    return dict(config=f"{ntm}x{npm1} uniform, {nr} realizations, {nsteps} steps of 1 h, "
                       f"checkpoint every {every}, RKG2 stages {int(s_r.min())}-{int(s_r.max())}",
                forward_s=round(fwd, 4), forward_ckpt_s=round(t1 - t0, 4),
                backward_s=round(t2 - t1, 4),
                adjoint_over_forward=round((t2 - t1) / fwd, 3),
                gradient_over_forward=round((t2 - t0) / fwd, 3),
                forward_ptl_s=round(fwd_ptl, 4),
                gradient_over_forward_ptl=round((t2 - t0) / fwd_ptl, 3),
                grad_norms=dict(f0=float(x.grad.norm()), eta=float(e.grad.norm()),
                                v0=float(v.grad.norm()), omega=float(o.grad.norm())))


def cfg5_bench(days=30, nr=128, ntm=512, npm1=1024, ckpt=24, adjoint_mem_gb=0):
    """Batched gradient assimilation at its workload: 512x1024, 128
    realizations, a window of `days` daily observations (truth run with eta =
    350 km^2/s, 5 G noise, mu^4 weights within 60 degrees of disk centre),
    Adam on eta, v0 and the initial maps.  One warm-up iteration on a one-day
    window (allocations), then one timed iteration of the full window: the
    forward pass with checkpoints, the adjoint pass and the optimiser step.
    Peak device memory = memory held after the adjoint pass (the library's
    buffers stay allocated) plus torch's peak above its current use."""
    import torch
    from hipft_amd.variational import synthetic_window, window_misfit
    torch.cuda.init()
    free_start, total = torch.cuda.mem_get_info()
    t0 = time.perf_counter()
    win = synthetic_window(ntm, npm1, nr, days)
    setup_s = time.perf_counter() - t0
    print(f"cfg5: window of {days} observations ready ({setup_s:.1f} s)", file=sys.stderr,
          flush=True)
    model, obs, wts, spo = win["model"], win["obs"], win["weights"], win["steps_per_obs"]
    if adjoint_mem_gb > 0:  # device memory the adjoint may use for stored RKG2 stages
        model.ft.set_tuning(adjoint_mem_mb=int(adjoint_mem_gb * 1024))

    def progress(msg):
        print(f"cfg5: {msg} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)

    le = torch.log(win["eta0"]).requires_grad_(True)
    lv = torch.log(win["v0_0"]).requires_grad_(True)
    om = win["omega0"]
    x0 = win["f0"].clone().requires_grad_(True)
    opt = torch.optim.Adam([{"params": [le, lv], "lr": 0.05}, {"params": [x0], "lr": 1.0}])

    def iteration(o, w):
        opt.zero_grad()
        torch.cuda.synchronize()
        a = time.perf_counter()
        J = window_misfit(model, x0, torch.exp(le), torch.exp(lv), om, o, w, spo,
                          checkpoint_every=ckpt, log=progress)
        total_J = J.sum()
        torch.cuda.synchronize()
        b = time.perf_counter()
        print(f"cfg5: forward over {len(o)} observations {b - a:.1f} s", file=sys.stderr,
              flush=True)
        total_J.backward()
        torch.cuda.synchronize()
        c = time.perf_counter()
        opt.step()
        torch.cuda.synchronize()
        return float(total_J.detach()), b - a, c - b, time.perf_counter() - c

    iteration(obs[:1], wts[:1])  # warm-up: allocations on a one-day window
    torch.cuda.reset_peak_memory_stats()
    J0, fwd, bwd, stp = iteration(obs, wts)
    torch.cuda.synchronize()
    held = free_start - torch.cuda.mem_get_info()[0]
    peak = held + torch.cuda.max_memory_allocated() - torch.cuda.memory_allocated()
    with torch.no_grad():
        J1 = float(window_misfit(model, x0, torch.exp(le), torch.exp(lv), om, obs, wts,
                                 spo, checkpoint_every=ckpt).sum())
    s_r = model.ft.diffusion_stats()[0]
    model_mem_mb = model.ft.get_tuning("adjoint_mem_mb")
    model.ft.close()
    it_s = fwd + bwd + stp
    steps = days * spo
    return dict(config=f"{ntm}x{npm1} clustered x2, {nr} realizations, {days} daily mu^4-"
                       f"weighted observations (5 G noise) of a truth run with eta = 350, Adam "
                       f"on eta, v0 and the initial maps, dt = 1 h, {steps} steps, checkpoints "
                       f"every {ckpt} steps, adjoint stage memory "
                       f"{model_mem_mb / 1024:.0f} GiB, RKG2 stages "
                       f"{int(s_r.min())}-{int(s_r.max())} (last step)",
                setup_s=round(setup_s, 2), forward_s=round(fwd, 3), adjoint_s=round(bwd, 3),
                optimizer_s=round(stp, 4), iteration_s=round(it_s, 3),
                adjoint_over_forward=round(bwd / fwd, 3),
                forward_realization_steps_per_s=round(nr * steps / fwd, 1),
                iterations_50_s=round(50 * it_s, 1),
                misfit_before=J0, misfit_after_one_step=J1,
                peak_device_GiB=round(peak / 2 ** 30, 2), device_GiB=round(total / 2 ** 30, 1))


def stream_ceiling(nbytes=2 << 30, reps=10):
    """Achievable HBM bandwidth of plain streaming kernels on this device
    (PyTorch's copy, add and addcmul over 2 GiB FP64 arrays: 1, 2 and 3
    reads per write), the practical ceiling beside the published 8 TB/s
    peak."""
    import torch
    n = nbytes // 8
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.rand(n, dtype=torch.float64, device="cuda")
    d = torch.rand(n, dtype=torch.float64, device="cuda")
    c = torch.empty_like(a)

    def rate(f, moved):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        return round(moved / ((time.perf_counter() - t0) / reps) / 1e9, 1)

    out = dict(copy_1r1w_GBs=rate(lambda: c.copy_(a), 2 * nbytes),
               add_2r1w_GBs=rate(lambda: torch.add(a, b, out=c), 3 * nbytes),
               addcmul_3r1w_GBs=rate(lambda: torch.addcmul(a, b, d, out=c), 4 * nbytes),
               fill_1w_GBs=rate(lambda: c.fill_(1.0), nbytes))
    del a, b, c, d
    torch.cuda.empty_cache()
    return out


def pmc_probe(m, nbytes=256 << 20):
    """Counter calibration copy of a known size, then one workload step.  The
    step is host-controlled, the launch sequence whose per-launch times the
    roofline's event pass measures."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    m.set_tuning(step_graph=0)
    # launch timing adds only event packets, no dispatches the counters see;
    # it gives the algorithmic bytes of this step's fused launches, the
    # reference for the measured bytes of the same launches
    m.set_timing(True)
    step(m)
    m.synchronize()
    ms, n, b = m.kernel_time("rkg2_fused")
    m.set_timing(False)
    print(json.dumps({"pmc_probe": "done", "copy_bytes": nbytes, "fused_launches": n,
                      "fused_alg_bytes_per_launch": round(b / max(n, 1))}), flush=True)


def load_traffic(kernel_sub):
    """HBM bytes per launch of a kernel, and of a whole workload step (every
    kernel of the step), from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size), and the
    algorithmic bytes per fused launch of the same probe step."""
    path = os.path.join(ROOT, "profiles", "step_pmc_batched64.json")
    if not os.path.exists(path):
        return None, None, None, None
    with open(path) as fh:
        d = json.load(fh)
    per, step_total, matched = None, 0.0, None
    for name, v in d.get("kernels", {}).items():
        step_total += v["hbm_bytes_per_launch"] * v["dispatches"]
        # "k_rkg2_fused2<4" matches the template's instantiations <4> and <4, 1024>
        if kernel_sub in name and name[name.index(kernel_sub) + len(kernel_sub)] in ">,(":
            per, matched = v["hbm_bytes_per_launch"], name
    alg = d.get("probe", {}).get("fused_alg_bytes_per_launch")
    return per, (step_total or None), matched, alg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nr", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the streaming-bandwidth ceiling (PyTorch copy/add)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint cost measurement (for profiling)")
    ap.add_argument("--retune", action="store_true", help="ignore the cached launch tuning")
    ap.add_argument("--no-step-graph", action="store_true",
                    help="host-controlled steps (no device-planned steps replayed as hipGraphs)")
    ap.add_argument("--cfg5", action="store_true",
                    help="only the batched gradient-assimilation measurement (512x1024, 128 "
                         "realizations, 30-day window; one Adam iteration, several minutes)")
    ap.add_argument("--cfg5-days", type=int, default=30)
    ap.add_argument("--cfg5-adjoint-mem-gb", type=float, default=210.0,
                    help="device memory for stored RKG2 stages in the cfg5 adjoint (0: the "
                         "library default, 32 GiB)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="for rocprofv3 --pmc runs: a 256 MiB device-to-device copy (counter "
                         "calibration) followed by one step of the workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; --gpus must be 1 "
                                   "(realizations would shard across GPUs with no per-step "
                                   "communication in independent mode)"}))
        sys.exit(2)
    if args.adjoint_only or not args.no_adjoint or not args.no_stream:
        # the adjoint section and the streaming ceiling share the device with
        # PyTorch: bring up torch's HIP context before libhft creates its own
        # streams
        import torch
        torch.cuda.init()
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_bench()}), flush=True)
        return
    if args.cfg5:
        print(json.dumps({"gradient_assimilation": cfg5_bench(days=args.cfg5_days,
                                                              adjoint_mem_gb=args.cfg5_adjoint_mem_gb)}), flush=True)
        return
    nr = args.nr
    wl = make_workload(nr)
    m = setup_model(wl)
    from hipft_amd import tune
    tuning = tune.autotune(m, use_cache=not args.retune,
                           log=(lambda msg: print(msg, file=sys.stderr, flush=True)))
    if args.no_step_graph:
        m.set_tuning(step_graph=0)
    if args.pmc_probe:
        pmc_probe(m)
        return
    for _ in range(args.warmup):
        step(m)
    # timed region: throughput of the plain step loop
    el = time_steps(m, args.steps)
    planned = dict(on=m.get_tuning("step_graph"), builds=m.get_tuning("step_graph_builds"),
                   replays=m.get_tuning("step_graph_hits"),
                   overflows=m.get_tuning("step_graph_overflows"))
    # second pass over as many steps with HIP events around every launch on
    # the library's stream (per-kernel times for the roofline; the events
    # themselves cost time, so this pass does not give the throughput)
    m.set_timing(True)
    el_ev = time_steps(m, args.steps)
    groups = {}
    for g in ("rkg2", "rkg2_fused", "weno", "ptl", "analysis"):
        ms, n, b = m.kernel_time(g)
        sb = m.kernel_stage_bytes(g)
        groups[g] = dict(ms=ms, launches=n, alg_GB=b / 1e9,
                         GBs=(b / 1e9) / (ms / 1e3) if ms > 0 else None)
        if sb != b:
            groups[g]["stage_equivalent_GB"] = sb / 1e9
    m.set_timing(False)
    s_last, cyc = m.diffusion_stats()
    value = nr * args.steps / el
    # dominant kernel: the multi-stage RKG2 launch when the tuning uses it
    fused = groups["rkg2_fused"]["launches"] > 0
    dom = "rkg2_fused" if fused else "rkg2"
    ms_d, n_d, b_d = groups[dom]["ms"], groups[dom]["launches"], groups[dom]["alg_GB"] * 1e9
    sb_d = groups[dom].get("stage_equivalent_GB", b_d / 1e9) * 1e9
    fk = {1: "k_rkg2_fused", 2: "k_rkg2_march", 3: "k_rkg2_fused2"}[
        tuning.get("fused_kernel", tune.DEFAULTS["fused_kernel"])]
    kname = f"{fk}<{tuning['fused_stages']}" if fused else "k_rkg2_stage"
    avg_ms = ms_d / max(n_d, 1)
    achieved = (b_d / 1e9) / (ms_d / 1e3) if ms_d > 0 else 0.0
    traffic, traffic_step, kfull, alg_probe = load_traffic(kname)
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                kernel=kfull or kname, avg_launch_ms=round(avg_ms, 5),
                alg_bytes_per_launch=round(b_d / max(n_d, 1)), launches=n_d,
                note="algorithmic bytes = the compulsory FP64 field passes of a launch: read "
                     "u_{k-1}, u_{k-2} and u0, write two stage levels, per realization it "
                     "advances (8 x 512 x 1023 bytes a pass); traffic = HBM bytes per launch "
                     "measured by rocprofv3 FETCH_SIZE + WRITE_SIZE (profiles/"
                     "step_pmc_batched64.json)")
    if fused and b_d > 0:
        # the same stages one per launch would move 5 passes each
        roof["stage_equivalent_GBs"] = round((sb_d / 1e9) / (ms_d / 1e3), 1)
        roof["temporal_blocking_gain"] = round(sb_d / b_d, 3)
    if traffic:
        # measured over algorithmic bytes of the same launches (the probe
        # step's); the timed steps' launch mix differs, so the measured rate
        # is the algorithmic one scaled by that ratio
        over = traffic / alg_probe if alg_probe else traffic / (b_d / max(n_d, 1))
        roof["traffic_over_alg"] = round(over, 3)
        roof["hbm_achieved"] = round(achieved * over, 1)
        roof["hbm_frac"] = round(roof["hbm_achieved"] / HBM_PEAK_GBS, 4)
    step_ms = 1e3 * el / args.steps
    if traffic_step:
        # whole step: measured DRAM bytes of one step (all kernels, the
        # committed counter profile) over the measured step time
        roof["step_hbm_GB"] = round(traffic_step / 1e9, 3)
        roof["step_hbm_frac"] = round(traffic_step / 1e9 / (step_ms / 1e3) / HBM_PEAK_GBS, 4)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "realization-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_ms, 3),
        "ms_per_step_with_kernel_events": round(1e3 * el_ev / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched: 512x1024 theta-clustered x2, {nr} realizations, FP64, "
                               "dt=1h Strang, WENO3/SSPRK(4,3) + RKG2/PTL, independent "
                               "per-realization stage counts, 200 seeded bipoles each",
                   "ntm": 512, "npm1": 1024, "realizations": nr,
                   "stages_last_step_min_max": [int(s_last.min()), int(s_last.max())],
                   "parallelism": "realization batch on 1 GPU",
                   "planned_steps": planned,
                   "launch_tuning": {k: tuning.get(k, tune.DEFAULTS[k]) for k in tune.KEYS}},
        "roofline": roof,
        "kernels": groups,
    }
    if not args.no_single:
        wl1 = {k: (v[:1] if isinstance(v, np.ndarray) and v.ndim >= 1 and len(v) == nr else v)
               for k, v in wl.items()}
        m1 = setup_model(wl1)
        tune.autotune(m1, use_cache=not args.retune)
        step(m1)
        k1 = max(4, args.steps // 2)
        el1 = time_steps(m1, k1)
        per_real = el1 / k1
        out["single"] = dict(value=round(k1 / el1, 3), unit="realization-steps/s",
                             ms_per_step=round(1e3 * per_real, 3),
                             batched_speedup_per_realization=round(value * per_real, 2),
                             note="one realization, launch shapes tuned for it (fused RKG2)")
        # the plain path: one realization, one RKG2 stage per launch, one
        # thread per point in the WENO3 stages
        m1.set_tuning(fused=0, advection_band=0, graphs=0)
        step(m1)
        el1p = time_steps(m1, k1)
        per_plain = el1p / k1
        out["single_plain"] = dict(value=round(k1 / el1p, 3), unit="realization-steps/s",
                                   ms_per_step=round(1e3 * per_plain, 3),
                                   batched_speedup_per_realization=round(value * per_plain, 2),
                                   note="one realization, one RKG2 stage per launch, pointwise "
                                        "WENO3 stages")
        m1.close()
    if not args.no_adjoint:
        m.close()
        out["adjoint"] = adjoint_bench()
    if not args.no_stream:
        sc = stream_ceiling()
        roof["stream_ceiling"] = sc
        if traffic:
            # the dominant kernel's measured HBM rate against the best of
            # the plain streaming kernels with reads and writes
            best = max(sc["copy_1r1w_GBs"], sc["add_2r1w_GBs"], sc["addcmul_3r1w_GBs"])
            roof["hbm_frac_of_stream_ceiling"] = round(roof["hbm_achieved"] / best, 4)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
