"""GPU parity: libhft (HIP, gfx950) against the CPU oracle and the reference
testsuite histories.

Operators and single steps are compared bit for bit with the oracle (same
inputs, same operation order, no FMA contraction, same polar-sum order).
Whole runs are compared with the reference's committed histories.
"""
import os

import numpy as np
import pytest

from conftest import golden
from cases import case_cfg, history_error

pytestmark = pytest.mark.gpu

from hipft_amd import workloads as W  # noqa: E402
from hipft_amd.core import FluxTransport  # noqa: E402
from hipft_amd import driver  # noqa: E402
from oracle.oracle import Oracle, run_case  # noqa: E402


def grid(ntm, npm1, clustered=True):
    t = W.clustered_theta(ntm, 2.0) if clustered else W.uniform_theta(ntm)
    return t, W.uniform_phi(npm1)


def random_state(o, seed, smooth=True):
    rng = np.random.default_rng(seed)
    t, p = o.mesh["t"], o.mesh["p"][: o.npm1]
    if smooth:
        m = W.bipole_map(t, p, 12, seed, fixed_sigma=None)
    else:
        m = rng.standard_normal((o.ntm, o.npm1))
    maps = np.stack([m * (1 + 0.1 * r) for r in range(o.nr)])
    o.ingest_pt(maps)
    return maps


def pair(ntm, npm1, nr, seed=1, clustered=True, mode="coupled", smooth=True):
    t, p = grid(ntm, npm1, clustered)
    o = Oracle(t, p, nr)
    maps = random_state(o, seed, smooth)
    g = FluxTransport(t, p, nr, mode=mode)
    g.set_state(maps, layout="pt")
    return o, g


def assert_same(a, b, what, rtol=0.0):
    if rtol == 0.0:
        bad = np.argwhere(a != b)
        assert bad.size == 0, f"{what}: {len(bad)} entries differ, first {bad[:3].tolist()}, " \
                              f"max |diff| {np.abs(a - b).max():.3e}"
    else:
        scale = np.abs(b).max()
        err = np.abs(a - b).max() / scale
        assert err <= rtol, f"{what}: max rel error {err:.3e} > {rtol:.1e}"


@pytest.mark.parametrize("shape", [(33, 65), (91, 181)])
def test_ingest_and_layout(gpu_available, shape):
    o, g = pair(*shape, nr=2)
    assert_same(g.get_state("tp"), o.get_f(), "ingested state")


@pytest.mark.parametrize("clustered", [True, False])
def test_diffusion_operator_bitexact(gpu_available, clustered):
    o, g = pair(47, 97, nr=3, clustered=clustered, smooth=False)
    eta = [100.0, 500.0, 1000.0]
    o.load_diffusion(eta)
    g.set_diffusion(eta)
    x = o.get_f()
    assert_same(g.apply_diffusion_operator(x), o.diffusion_operator(x), "L x")
    dt, dtr = g.dt_euler()
    assert dt == o.dt_euler
    assert all(dtr[r] == o.dt_euler_r(r) for r in range(3))


def _flows(o, g, nr, with_profiles=True):
    t = o.mesh["t"]
    p0, p2, p4 = (np.array([c * (1.0 + 0.05 * r) for r in range(nr)]) for c in W.dr_coefficients())
    vt_prof = np.stack([W.meridional_profile(t, 10.0 + 5.0 * r) for r in range(nr)])
    o.flow_reset()
    if with_profiles:
        vt = np.ascontiguousarray(np.broadcast_to(vt_prof[:, None, :], (nr, o.npm, o.nt)))
        o.flow_set(vt, np.zeros((nr, o.np_, o.ntm)))
        g.set_flow_profiles(vt_prof, None)
    o.flow_add_dr(p0, p2, p4)
    o.set_lf_alpha()
    g.set_flow_analytic(p0=p0, p2=p2, p4=p4)


def test_advection_operator_bitexact(gpu_available):
    o, g = pair(61, 121, nr=2, smooth=True)
    _flows(o, g, 2)
    x = o.get_f()
    assert_same(g.apply_advection_operator(x), o.advection_operator(x), "A x")


def test_advection_operator_mf_rigid_const(gpu_available):
    o, g = pair(45, 91, nr=2, clustered=False)
    p1, p3, p5 = [22.0, 17.8], [11.0, 9.9], [-28.0, -30.8]
    o.flow_reset()
    o.flow_add_mf(p1, p3, p5)
    o.flow_add_rigid(0.3)
    o.flow_add_vt_const(-0.05)
    o.set_lf_alpha()
    g.set_flow_analytic(p1=p1, p3=p3, p5=p5, vp_rigid_omega=0.3, vt_const=-0.05)
    x = o.get_f()
    assert_same(g.apply_advection_operator(x), o.advection_operator(x), "A x (MF+rigid+vt)")
    for strang in (0, 1):
        assert g.flow_dtmax(strang)[0] == o.flow_dtmax(strang)


def test_advection_step_bitexact(gpu_available):
    o, g = pair(61, 121, nr=2)
    _flows(o, g, 2)
    for _ in range(3):
        o.advection_step(0.5)
        g.advection_step(0.5)
    assert_same(g.get_state("tp"), o.get_f(), "3 SSPRK(4,3) steps")


@pytest.mark.parametrize("band", [0, 1])
def test_independent_advection_cfl_substeps(gpu_available, band):
    """Independent mode, an advection step longer than some realizations' flow
    CFL limits: realization r takes n_r = ceil(dt / limit_r) SSPRK(4,3) steps
    of dt / n_r (limit_r = the single-realization oracle's flow_dtmax(0)),
    bit for bit; with the sub-steps switched off every realization takes one
    step of dt."""
    t, p = grid(61, 121)
    nr = 4
    scale = [1.0, 2.2, 3.5, 1.4]  # differential rotation sets the limits here
    vt = np.stack([W.meridional_profile(t, 10.0 + 4.0 * r) for r in range(nr)])
    p0, p2, p4 = (np.array([c * scale[r] for r in range(nr)]) for c in W.dr_coefficients())
    singles, maps = [], []
    for r in range(nr):
        o = Oracle(t, p, 1)
        maps.append(random_state(o, 30 + r)[0])
        o.flow_reset()
        o.flow_set(np.ascontiguousarray(np.broadcast_to(vt[r][None, None, :], (1, o.npm, o.nt))),
                   np.zeros((1, o.np_, o.ntm)))
        o.flow_add_dr([p0[r]], [p2[r]], [p4[r]])
        o.set_lf_alpha()
        singles.append(o)
    lim = np.array([o.flow_dtmax(0) for o in singles])
    dt = 0.98 * lim.max()
    n = np.maximum(1, np.ceil(dt / lim)).astype(int)
    assert n.min() == 1 and len(set(n.tolist())) == 4, n

    def gpu(substeps):
        g = FluxTransport(t, p, nr, mode="independent")
        g.set_state(np.stack(maps), layout="pt")
        g.set_flow_profiles(vt, None)
        g.set_flow_analytic(p0=p0, p2=p2, p4=p4)
        g.set_tuning(advection_band=band, advection_substeps=substeps)
        assert np.array_equal(g.flow_dtmax(0)[1], lim)
        for _ in range(2):
            g.advection_step(dt)
        f = g.get_state("tp")
        g.close()
        return f

    f = gpu(1)
    for r, o in enumerate(singles):
        for _ in range(2 * n[r]):
            o.advection_step(dt / n[r])
        assert_same(f[r:r + 1], o.get_f(), f"realization {r} with {n[r]} sub-steps")
    f0 = gpu(0)
    for r in range(nr):
        o = Oracle(t, p, 1)
        o.ingest_pt(maps[r][None])
        o.flow_reset()
        o.flow_set(np.ascontiguousarray(np.broadcast_to(vt[r][None, None, :], (1, o.npm, o.nt))),
                   np.zeros((1, o.np_, o.ntm)))
        o.flow_add_dr([p0[r]], [p2[r]], [p4[r]])
        o.set_lf_alpha()
        for _ in range(2):
            o.advection_step(dt)
        assert_same(f0[r:r + 1], o.get_f(), f"realization {r} without sub-steps")


def test_attenuated_flow_advection(gpu_available):
    o, g = pair(45, 91, nr=2)
    p0, p2, p4 = W.dr_coefficients()
    att = [300.0, 150.0]
    o.flow_reset()
    o.flow_add_mf([22.0, 20.0], [11.0, 11.0], [-28.0, -28.0])
    o.flow_add_dr([p0] * 2, [p2] * 2, [p4] * 2)
    o.flow_attenuate(att)
    o.set_lf_alpha()
    g.set_flow_analytic(p0=p0, p2=p2, p4=p4, p1=[22.0, 20.0], p3=11.0, p5=-28.0)
    g.set_flow_attenuation(att)
    x = o.get_f()
    # tanh of the device math library may differ from glibc in the last bit
    assert_same(g.apply_advection_operator(x), o.advection_operator(x), "A x (attenuated)", 1e-12)
    dt_o = o.flow_dtmax(0)
    dt_g = g.flow_dtmax(0)[0]
    assert abs(dt_g - dt_o) <= 1e-13 * dt_o


def _flow_2d(o, nrf, seed):
    """phi-dependent flows in code units, periodic in the reference's seam
    columns: vt [nrf][npm][nt], vp [nrf][np][ntm]."""
    rng = np.random.default_rng(seed)
    t, p = o.mesh["t"][: o.ntm], o.mesh["p"]
    vt = np.empty((nrf, o.npm, o.nt))
    vp = np.empty((nrf, o.np_, o.ntm))
    for r in range(nrf):
        a, b, ph = rng.uniform(0.5, 1.5), rng.uniform(0.1, 0.4), rng.uniform(0, 2 * np.pi)
        mf = W.meridional_profile(t, 15.0 * a)  # on the nt theta faces
        vt[r] = mf[None, :] * (1.0 + b * np.cos(p[: o.npm, None] + ph))
        vp[r] = 2e-4 * np.sin(t)[None, :] * np.sin(2 * p[: o.np_, None] + ph)
    for v in (vt, vp):
        v[:, 0] = v[:, -2]
        v[:, -1] = v[:, 1]
    return vt, vp


@pytest.mark.parametrize("shared", [True, False])
def test_flows_from_arrays_bitexact(gpu_available, shared):
    """2-D base flows (flows from files) plus analytic meridional flow and
    differential rotation: flows, Lax-Friedrichs alphas, the WENO3 operator,
    the CFL limit and SSPRK(4,3) steps match the oracle bit for bit."""
    nr = 3
    o, g = pair(45, 91, nr=nr)
    vt, vp = _flow_2d(o, 1 if shared else nr, 7)
    p0, p2, p4 = (np.array([c * (1.0 + 0.1 * r) for r in range(nr)]) for c in W.dr_coefficients())
    p1 = [22.0, 18.0, 26.0]
    o.flow_reset()
    o.flow_set(np.broadcast_to(vt, (nr,) + vt.shape[1:]), np.broadcast_to(vp, (nr,) + vp.shape[1:]))
    o.flow_add_mf(p1, [11.0] * nr, [-28.0] * nr)
    o.flow_add_dr(p0, p2, p4)
    o.flow_add_rigid(0.2)
    o.flow_add_vt_const(-0.03)
    o.set_lf_alpha()
    g.set_flow_analytic(p0=p0, p2=p2, p4=p4, p1=p1, p3=11.0, p5=-28.0, vp_rigid_omega=0.2,
                        vt_const=-0.03)
    g.set_flow_arrays(vt, vp)
    x = o.get_f()
    assert_same(g.apply_advection_operator(x), o.advection_operator(x), "A x (2-D flows)")
    dt_o = o.flow_dtmax(0)
    assert g.flow_dtmax(0)[0] == dt_o
    for _ in range(2):
        o.advection_step(0.4 * dt_o)
        g.advection_step(0.4 * dt_o)
    assert_same(g.get_state("tp"), o.get_f(), "SSPRK(4,3) with 2-D flows")
    # removing the base flows leaves the analytic flows
    g.set_flow_arrays(None, None)
    o.flow_reset()
    o.flow_add_mf(p1, [11.0] * nr, [-28.0] * nr)
    o.flow_add_dr(p0, p2, p4)
    o.flow_add_rigid(0.2)
    o.flow_add_vt_const(-0.03)
    o.set_lf_alpha()
    x = o.get_f()
    assert_same(g.apply_advection_operator(x), o.advection_operator(x), "A x (base removed)")


def test_flows_from_arrays_attenuated(gpu_available):
    """2-D base flows attenuated by |Br| every step (update_flow with
    flow_attenuate): within 1e-12 of the oracle (device tanh)."""
    nr = 2
    o, g = pair(45, 91, nr=nr)
    vt, vp = _flow_2d(o, 1, 11)
    p0, p2, p4 = W.dr_coefficients()
    att = [250.0, 120.0]

    def oracle_flow():
        o.flow_reset()
        o.flow_set(np.broadcast_to(vt, (nr,) + vt.shape[1:]), np.broadcast_to(vp, (nr,) + vp.shape[1:]))
        o.flow_add_dr([p0] * nr, [p2] * nr, [p4] * nr)
        o.flow_attenuate(att)
        o.set_lf_alpha()

    g.set_flow_arrays(vt, vp)
    g.set_flow_analytic(p0=p0, p2=p2, p4=p4)
    g.set_flow_attenuation(att)
    oracle_flow()
    x = o.get_f()
    assert_same(g.apply_advection_operator(x), o.advection_operator(x), "A x (2-D, attenuated)",
                1e-12)
    for _ in range(3):
        o.advection_step(2.0)
        g.advection_step(2.0)
        oracle_flow()
        g.update_flow()
    assert_same(g.get_state("tp"), o.get_f(), "advection with attenuated 2-D flows", 1e-11)


def test_rkg2_fixed_cycles_bitexact(gpu_available):
    o, g = pair(47, 97, nr=2)
    eta = [300.0, 700.0]
    o.load_diffusion(eta)
    g.set_diffusion(eta)
    res = o.diffusion_step(2.0, subcycles=2)
    g.diffusion_step(2.0, subcycles=2)
    s, cyc = g.diffusion_stats()
    assert list(s) == [res["s"][-1]] * 2 and list(cyc) == [2, 2]
    assert_same(g.get_state("tp"), o.get_f(), "RKG2 2 cycles")


def test_ptl_diffusion_step_bitexact(gpu_available):
    o, g = pair(91, 181, nr=3)
    eta = [100.0, 400.0, 900.0]
    o.load_diffusion(eta)
    g.set_diffusion(eta)
    recs = []
    res = o.diffusion_step(24.0, subcycles=0)
    g.diffusion_step(24.0, subcycles=0, callback=lambda cyc, rows: recs.append(rows))
    assert len(recs) == res["ncycles"]
    for k, rows in enumerate(recs):
        assert rows[0][0] == res["time"][k]
        assert rows[0][1] == res["dt"][k]
        assert rows[0][2] == res["s"][k]
    assert_same(g.get_state("tp"), o.get_f(), "RKG2 + PTL")


@pytest.mark.parametrize("mode", ["coupled", "independent"])
def test_ptl_ties_bitexact(gpu_available, mode):
    """The PTL step limit where max |L f| is reached at several points: a
    map made of one 4-column pattern repeated exactly along phi (512
    columns), so |L f| ties at 11 points of the equator row, several in one
    column tile (the Laplacian pass records a tie count above one) and in
    both tiles of the row; the limit binds (two cycles in 24 h).  Cycle
    lengths, stage counts and the state equal the oracle's bit for bit."""
    t, p = W.uniform_theta(91), W.uniform_phi(513)
    th = np.asarray(t)[:, None]
    per = 4
    block = np.sin(th) ** 2 * np.cos(2 * np.pi * np.arange(per)[None, :] / per) * 10.0 \
        + np.sin(th) ** 4 * 3.0
    m = np.tile(block, (1, 512 // per))
    m = np.concatenate([m, m[:, :1]], axis=1)
    maps = np.stack([m, 0.5 * m])
    eta = [800.0, 800.0]
    g = FluxTransport(t, p, 2, mode=mode)
    g.set_state(maps, layout="pt")
    g.set_diffusion(eta)
    if mode == "coupled":
        o = Oracle(t, p, 2)
        o.ingest_pt(maps)
        o.load_diffusion(eta)
        recs = []
        res = o.diffusion_step(24.0, subcycles=0)
        g.diffusion_step(24.0, subcycles=0, callback=lambda cyc, rows: recs.append(rows))
        assert res["ncycles"] >= 2, res["ncycles"]
        assert len(recs) == res["ncycles"]
        for k, rows in enumerate(recs):
            assert rows[0][1] == res["dt"][k]
            assert rows[0][2] == res["s"][k]
        assert_same(g.get_state("tp"), o.get_f(), "PTL with ties")
    else:
        g.diffusion_step(24.0, subcycles=0)
        f = g.get_state("tp")
        _, cyc = g.diffusion_stats()
        assert cyc.max() >= 2, cyc
        for r in range(2):
            o1 = Oracle(t, p, 1)
            o1.ingest_pt(maps[r:r + 1])
            o1.load_diffusion([eta[r]])
            o1.diffusion_step(24.0, subcycles=0)
            assert_same(f[r:r + 1], o1.get_f(), f"PTL with ties, realization {r}")
    g.close()


def test_independent_mode_matches_single_runs(gpu_available):
    t, p = grid(61, 121)
    eta = [100.0, 350.0, 1000.0]
    g = FluxTransport(t, p, 3, mode="independent")
    singles = []
    maps = []
    for r in range(3):
        o = Oracle(t, p, 1)
        m = random_state(o, 10 + r)
        maps.append(m[0])
        o.load_diffusion([eta[r]])
        singles.append(o)
    g.set_state(np.stack(maps), layout="pt")
    g.set_diffusion(eta)
    out = []
    for r, o in enumerate(singles):
        out.append(o.diffusion_step(12.0, subcycles=0))
    g.diffusion_step(12.0, subcycles=0)
    s, cyc = g.diffusion_stats()
    f = g.get_state("tp")
    for r, o in enumerate(singles):
        assert s[r] == out[r]["s"][-1] and cyc[r] == out[r]["ncycles"]
        assert_same(f[r:r + 1], o.get_f(), f"independent realization {r}")
    assert len(set(s.tolist())) > 1  # stage counts differ per realization


@pytest.mark.parametrize("nr", [2, 5])
def test_analysis_matches_oracle(gpu_available, nr):
    # nr = 5: the final reduction's 16 lanes per realization span two wavefronts
    o, g = pair(91, 181, nr=nr)
    ho = o.analysis()
    hg = g.analysis()
    assert np.array_equal(hg[:, :3], ho[:, :3])
    scale = np.maximum(np.abs(ho[:, 3:5]).max(), 1.0)
    assert np.abs(hg[:, 3:11] - ho[:, 3:11]).max() <= 1e-13 * scale
    assert np.allclose(hg[:, 11:], ho[:, 11:], rtol=1e-12, atol=1e-14 * np.abs(ho[:, 3]).max() / 4.84416e21)


@pytest.mark.parametrize("add_mode,balance,lat", [(False, False, 0.0), (True, False, 0.0),
                                                  (False, True, 0.0), (False, False, 60.0)])
def test_assimilation_matches_oracle(gpu_available, add_mode, balance, lat):
    o, g = pair(61, 121, nr=2)
    t, p = o.mesh["t"], o.mesh["p"][: o.npm1]
    rng = np.random.default_rng(5)
    data = W.bipole_map(t, p, 8, 99) + rng.normal(0, 5.0, (o.ntm, o.npm1))
    mu = W.mu_map(t, p, 120.0)
    d3 = np.stack([data, np.zeros_like(data), mu])
    d4 = np.ascontiguousarray(np.broadcast_to(d3, (2,) + d3.shape))
    kw = dict(custom_from_mu=True, mu_power=[4.0, 2.0], mu_limit=[0.5, 0.3], lat_limit=lat,
              mult_fac=1.1, add_mode=add_mode, balance=balance)
    o.assimilate(d4, **kw)
    g.assimilate(d3, **kw)
    assert_same(g.get_state("tp"), o.get_f(), "assimilation", 1e-14)


@pytest.mark.parametrize("power,limit,mult", [(4.0, 0.5, 1.0), (3.0, 0.5, 1.3)])
def test_assimilation_closed_form(gpu_available, power, limit, mult):
    """The GPU blend against the closed form f + mu^p (mult d - f), strict
    cut mu > limit (mu exactly at the limit keeps the state), the seam column
    taking its periodic twin's increment (data column 1 unused) and pole rows
    left as blended; the same inputs as the oracle's closed-form pins."""
    from test_oracle_math import _mu_field, assim_expected
    o, g = pair(31, 61, nr=1, clustered=False)
    t, p = o.mesh["t"], o.mesh["p"][: o.npm1]
    f0 = g.get_state("tp")[0].copy()
    data = 40.0 * W.bipole_map(t, p, 5, 77) + 7.0
    data[:, 0] = 1.0e4  # a first meridian that disagrees with the last: never read
    mu = _mu_field(o, t, p)
    g.assimilate(np.stack([data, np.zeros_like(data), mu]), custom_from_mu=True,
                 mu_power=power, mu_limit=limit, mult_fac=mult)
    f = g.get_state("tp")[0]
    exp = assim_expected(f0, data, mu, power, limit, mult)
    assert np.abs(f - exp).max() <= 4e-15 * np.abs(exp).max()
    assert np.array_equal(f[5, 1:-1], f0[5, 1:-1])
    assert np.all(f[6, 1:-1] != f0[6, 1:-1])
    assert np.ptp(f[1:-1, 0]) > 0.5 and np.ptp(f[1:-1, -1]) > 0.5


@pytest.mark.parametrize("case", ["advect_gaussians_phi", "advect_gaussians_theta",
                                  "advect_gaussians_phi_theta", "diffuse_soccer",
                                  "diffuse_advect_soccer"])
def test_testsuite_case_vs_reference(gpu_available, case):
    cfg = case_cfg(case)
    sol, num, R = driver.run(cfg)
    ref = golden(case)
    assert sol.shape == ref["sol"].shape
    err, verr, _ = history_error(sol, ref["sol"])
    assert err <= 5e-12, f"{case}: history error {err:.3e}"
    assert verr <= 1e-6
    assert np.array_equal(num[..., 6:8], ref["num"][..., 6:8])
    assert np.allclose(num[..., :6], ref["num"][..., :6], rtol=1e-14, atol=0)


@pytest.mark.parametrize("case", ["diffuse_soccer", "advect_gaussians_phi_theta"])
def test_testsuite_prefix_vs_oracle(gpu_available, case):
    cfg = case_cfg(case)
    sol_g, num_g, R = driver.run(cfg, max_steps=12)
    sol_o, num_o, o = run_case(cfg, max_steps=12)
    assert np.array_equal(num_g, num_o)
    assert_same(R.model.get_state("tp"), o.get_f(), f"{case} state after 12 steps")
    err, _, _ = history_error(sol_g, sol_o)
    assert err <= 1e-13


def test_diffuse_dipole_vs_oracle_and_reference(gpu_available):
    """512x1024, one 8760 h step: 7 PTL cycles, 36,205 RKG2 stages.

    Against the oracle (tests/golden/oracle_diffuse_dipole.npz): cycle times,
    stage counts and cycle counts identical, histories to summation order.
    Against the reference fixture: stage and cycle counts identical, the
    first three cycle times to 5e-12; from the 4th PTL cycle on the cycle
    length differs by ~2e-8 relative because that cycle's PTL step amplifies
    last-bit differences of the state ~1e4-fold (one unit in the last place
    of the input map moves it by 2-4e-8; test_diffuse_dipole_ptl_conditioning
    shows this on the CPU), so the remaining rows are held to 1e-7.
    """
    case = "diffuse_dipole"
    sol, num, R = driver.run(case_cfg(case))
    orc = golden("oracle_diffuse_dipole")
    assert np.array_equal(num, orc["num"])
    err, _, cols = history_error(sol, orc["sol"])
    assert err <= 1e-13, cols
    ref = golden(case)
    assert np.array_equal(num[..., 6:8], ref["num"][..., 6:8])
    # rows: initial, cycles 1..7, final; cycles 1-3 before the ill-conditioned 4th
    assert np.allclose(num[:, 1:4, 1], ref["num"][:, 1:4, 1], rtol=5e-12, atol=0)
    assert np.allclose(num[..., 1], ref["num"][..., 1], rtol=1e-7)
    err, _, cols = history_error(sol[:, :4], ref["sol"][:, :4])
    assert err <= 1e-11, cols
    err, _, cols = history_error(sol, ref["sol"])
    assert err <= 1e-7, cols


def test_atten_map_vs_reference(gpu_available):
    case = "diffuse_advect_atten_map_1cr"
    sol, num, R = driver.run(case_cfg(case))
    ref = golden(case)
    assert sol.shape == ref["sol"].shape
    assert np.array_equal(num[..., 6:8], ref["num"][..., 6:8])
    err, _, cols = history_error(sol, ref["sol"])
    assert err <= 1e-10, cols


@pytest.mark.parametrize("kernel", [1, 2, 3])
@pytest.mark.parametrize("T,H,B", [(2, 4, 0), (3, 7, 0), (4, 32, 0), (4, 5, 0), (4, 32, 40),
                                   (3, 64, 1000), (4, 2, 0), (3, 3, 0)])
def test_fused_rkg2_bitexact(gpu_available, T, H, B, kernel):
    """Several RKG2 stages per launch (LDS row rings, band halos) give the
    oracle's result bit for bit, across band boundaries and the poles, also
    when the bands thin out as realizations finish (B = target blocks)."""
    o, g = pair(91, 181, nr=3)
    eta = [150.0, 450.0, 950.0]
    o.load_diffusion(eta)
    g.set_diffusion(eta)
    g.set_tuning(fused=1, fused_stages=T, band_rows=H, fused_min_realizations=1,
                 fused_target_blocks=B, fused_kernel=kernel)
    o.diffusion_step(6.0, subcycles=0)
    g.diffusion_step(6.0, subcycles=0)
    assert_same(g.get_state("tp"), o.get_f(), f"fused kernel {kernel} T={T} H={H} B={B}")


@pytest.mark.parametrize("H,B", [(5, 0), (32, 0), (64, 0), (32, 40)])
def test_fused_rkg2_five_stages_bitexact(gpu_available, H, B):
    """Five RKG2 stages per launch (two-column kernel; the fifth stage takes
    u0 and dty0 of its row from the slot the step reuses) give the oracle's
    result bit for bit."""
    o, g = pair(91, 181, nr=3)
    eta = [150.0, 450.0, 950.0]
    o.load_diffusion(eta)
    g.set_diffusion(eta)
    g.set_tuning(fused=1, fused_kernel=3, fused_stages=5, band_rows=H, fused_min_realizations=1,
                 fused_target_blocks=B)
    o.diffusion_step(6.0, subcycles=0)
    g.diffusion_step(6.0, subcycles=0)
    assert_same(g.get_state("tp"), o.get_f(), f"fused kernel 3 T=5 H={H} B={B}")


def test_fused_five_stages_full_size_bitexact(gpu_available):
    """Five stages per launch on the 512x1024 batch with fixed 32- and
    64-row bands (general march steps: interior steps are off for five
    stages) against one stage per launch, bit for bit."""
    t, p = grid(512, 1024)
    nr = 6
    maps = np.stack([W.bipole_map(t, p, 40, 300 + r) for r in range(nr)])
    eta = np.array([100.0, 180.0, 320.0, 560.0, 800.0, 1000.0])
    out = {}
    for key, tun in (("one-stage", dict(fused=0)),
                     ("H=32", dict(fused=1, fused_stages=5, band_rows=32)),
                     ("H=64", dict(fused=1, fused_stages=5, band_rows=64))):
        m = FluxTransport(t, p, nr, mode="independent")
        m.set_state(maps, layout="pt")
        m.set_diffusion(eta)
        m.set_tuning(fused_min_realizations=1, fused_kernel=3, fused_bands=0,
                     fused_target_blocks=0, fused_body=1, **tun)
        m.diffusion_step(1.0, subcycles=0)
        out[key] = m.get_state("tp")
        m.close()
    for key in ("H=32", "H=64"):
        assert_same(out[key], out["one-stage"], f"512x1024 five stages {key}")


@pytest.mark.parametrize("T", [3, 4, 5])
def test_fused_round_aware_bands_bitexact(gpu_available, T):
    """Band counts chosen per launch against whole rounds of co-resident
    blocks (fused_bands=1): a 181-row grid (ragged last band, independent
    stage counts) against the oracle, and the 512x1024 batch against the
    one-stage-per-launch path, bit for bit."""
    o, g = pair(181, 240, nr=5, mode="independent")
    maps0 = g.get_state("pt")
    eta = [120.0, 300.0, 520.0, 760.0, 990.0]
    g.set_diffusion(eta)
    g.set_tuning(fused=1, fused_stages=T, fused_min_realizations=1, fused_kernel=3,
                 fused_bands=1)
    g.diffusion_step(6.0, subcycles=0)
    f = g.get_state("tp")
    for r in range(5):
        o1 = Oracle(o.mesh["t"], o.mesh["p"][: o.npm1], 1)
        o1.ingest_pt(maps0[r:r + 1])
        o1.load_diffusion([eta[r]])
        o1.diffusion_step(6.0, subcycles=0)
        assert_same(f[r:r + 1], o1.get_f(), f"T={T} realization {r}")
    t, p = grid(512, 1024)
    nr = 6
    maps = np.stack([W.bipole_map(t, p, 40, 200 + r) for r in range(nr)])
    out = []
    for fused in (0, 1):
        m = FluxTransport(t, p, nr, mode="independent")
        m.set_state(maps, layout="pt")
        m.set_diffusion(np.array([100.0, 180.0, 320.0, 560.0, 800.0, 1000.0]))
        m.set_tuning(fused=fused, fused_stages=T, fused_min_realizations=1, fused_kernel=3,
                     fused_bands=fused)
        m.diffusion_step(1.0, subcycles=0)
        out.append(m.get_state("tp"))
        m.close()
    assert_same(out[1], out[0], f"512x1024 round-aware bands T={T}")
    g.close()


@pytest.mark.parametrize("T", [2, 3, 4])
def test_fused_interior_steps_bitexact(gpu_available, T):
    """The two-column kernel's interior march steps (no row tests, buffer
    loads and stores, compile-time ring slots) against the oracle on a
    91x181 grid with bands tall enough for them (runtime LDS row length),
    and against the same kernel without them on the 512x1024 batch (LDS row
    length 1024), bit for bit.  Five stages per launch run the general
    steps only."""
    o, g = pair(91, 181, nr=3)
    eta = [150.0, 450.0, 950.0]
    o.load_diffusion(eta)
    g.set_diffusion(eta)
    g.set_tuning(fused=1, fused_stages=T, band_rows=40, fused_min_realizations=1,
                 fused_kernel=3, fused_body=1)
    o.diffusion_step(6.0, subcycles=0)
    g.diffusion_step(6.0, subcycles=0)
    assert_same(g.get_state("tp"), o.get_f(), f"interior steps T={T} 91x181")
    g.close()
    t, p = grid(512, 1024)
    nr = 6
    maps = np.stack([W.bipole_map(t, p, 40, 300 + r) for r in range(nr)])
    for shape in (dict(fused_bands=1), dict(fused_bands=0, fused_target_blocks=0, band_rows=64)):
        out = []
        for body in (0, 1):
            m = FluxTransport(t, p, nr, mode="independent")
            m.set_state(maps, layout="pt")
            m.set_diffusion(np.array([100.0, 180.0, 320.0, 560.0, 800.0, 1000.0]))
            m.set_tuning(fused=1, fused_stages=T, fused_min_realizations=1, fused_kernel=3,
                         fused_body=body, **shape)
            m.diffusion_step(1.0, subcycles=0)
            out.append(m.get_state("tp"))
            m.close()
        assert_same(out[1], out[0], f"512x1024 interior steps T={T} {shape}")


@pytest.mark.parametrize("kernel", [1, 2, 3])
def test_fused_matches_single_stage_full_size(gpu_available, kernel):
    """512x1024 batch, independent stage counts: fused and one-stage-per-launch
    paths agree bit for bit."""
    t, p = grid(512, 1024)
    nr = 6
    maps = np.stack([W.bipole_map(t, p, 40, 100 + r) for r in range(nr)])
    eta = np.array([100.0, 180.0, 320.0, 560.0, 800.0, 1000.0])
    out = []
    for fused in (0, 1):
        g = FluxTransport(t, p, nr, mode="independent")
        g.set_state(maps, layout="pt")
        g.set_diffusion(eta)
        g.set_tuning(fused=fused, fused_min_realizations=1, fused_target_blocks=512 * fused,
                     fused_kernel=kernel)
        g.diffusion_step(1.0, subcycles=0)
        out.append(g.get_state("tp"))
        s, _ = g.diffusion_stats()
        g.close()
    assert len(set(s.tolist())) == nr
    assert_same(out[1], out[0], "fused vs single-stage at 512x1024")


@pytest.mark.parametrize("H", [1, 2, 5, 16, 61])
def test_banded_advection_bitexact(gpu_available, H):
    """Banded WENO3 stages (faces computed once, bands marching in theta) give
    the oracle's operator and SSPRK(4,3) steps bit for bit for every band
    height, including bands that begin or end on a pole row."""
    o, g = pair(61, 121, nr=2)
    _flows(o, g, 2)
    g.set_tuning(advection_band=1, advection_band_rows=H)
    x = o.get_f()
    assert_same(g.apply_advection_operator(x), o.advection_operator(x), f"A x band H={H}")
    for _ in range(2):
        o.advection_step(0.5)
        g.advection_step(0.5)
    assert_same(g.get_state("tp"), o.get_f(), f"SSPRK(4,3) band H={H}")


@pytest.mark.parametrize("band", [0, 1])
def test_weno_short_divides_and_their_fallback(gpu_available, band):
    """WENO3 weights: the short divides (shared refined reciprocal, no
    scaling or fix-up steps) where the smoothness terms lie in [eps, 2^100],
    the full divide elsewhere.  A map with spikes of 1e40 (smoothness terms
    far above 2^100 next to ordinary ones, so lanes of one wavefront take
    both paths) and a smooth map give the oracle's operator and SSPRK(4,3)
    steps bit for bit, with the short divides on and off."""
    for spikes in (False, True):
        o, g = pair(61, 121, nr=2, seed=5)
        if spikes:
            maps = g.get_state("pt").copy()
            rng = np.random.default_rng(7)
            idx = rng.integers(0, maps.size, 40)
            maps.reshape(-1)[idx] = 1e40 * rng.standard_normal(40)
            o.ingest_pt(maps)
            g.set_state(maps, layout="pt")
            assert_same(g.get_state("tp"), o.get_f(), "spiked state")
        _flows(o, g, 2)
        for fast in (1, 0):
            g.set_tuning(advection_band=band, advection_band_rows=16, weno_fastdiv=fast)
            assert g.get_tuning("weno_fastdiv") == fast
            x = o.get_f()
            assert_same(g.apply_advection_operator(x), o.advection_operator(x),
                        f"A x band={band} spikes={spikes} fastdiv={fast}")
        o.advection_step(0.5)
        g.advection_step(0.5)
        assert_same(g.get_state("tp"), o.get_f(), f"SSPRK(4,3) band={band} spikes={spikes}")
        g.close()


@pytest.mark.parametrize("H", [64, 128])
def test_banded_advection_tall_bands_512_rows(gpu_available, H):
    """512x1024 clustered, bands of 64 and 128 theta rows (the tuned shapes
    of the benchmark batch): operator and two SSPRK(4,3) steps equal the
    oracle bit for bit."""
    o, g = pair(512, 1024, nr=2)
    _flows(o, g, 2)
    g.set_tuning(advection_band=1, advection_band_rows=H)
    x = o.get_f()
    assert_same(g.apply_advection_operator(x), o.advection_operator(x), f"A x band H={H}")
    for _ in range(2):
        o.advection_step(0.5)
        g.advection_step(0.5)
    assert_same(g.get_state("tp"), o.get_f(), f"SSPRK(4,3) band H={H} at 512x1024")


@pytest.mark.parametrize("atten", [False, True])
def test_banded_advection_matches_pointwise_full_size(gpu_available, atten):
    """512x1024, 1-D flow profiles and attenuated (2-D) flows: banded and
    one-thread-per-point WENO3 stages agree bit for bit."""
    t, p = grid(512, 1024)
    nr = 3
    maps = np.stack([W.bipole_map(t, p, 40, 200 + r) for r in range(nr)])
    vt = np.stack([W.meridional_profile(t, 12.0 + 3.0 * r) for r in range(nr)])
    p0, p2, p4 = (np.array([c * (1.0 + 0.04 * r) for r in range(nr)]) for c in W.dr_coefficients())
    out = []
    for band in (0, 1):
        g = FluxTransport(t, p, nr, mode="independent")
        g.set_state(maps, layout="pt")
        g.set_flow_profiles(vt, None)
        g.set_flow_analytic(p0=p0, p2=p2, p4=p4)
        if atten:
            g.set_flow_attenuation([300.0, 200.0, 150.0])
            g.update_flow()
        g.set_tuning(advection_band=band)
        g.advection_step(0.5)
        out.append(g.get_state("tp"))
        g.close()
    assert_same(out[1], out[0], "banded vs pointwise WENO3 at 512x1024")


@pytest.mark.parametrize("fused", [0, 1])
def test_graph_replay_matches_direct_launches(gpu_available, fused):
    """RKG2 super-steps replayed from captured hipGraphs give the same states
    bit for bit as direct launches, over Strang steps with fixed sub-cycles
    (stage counts repeat, so graphs are reused) and with PTL sub-cycling."""
    t, p = grid(91, 181)
    nr = 3
    maps = np.stack([W.bipole_map(t, p, 20, 300 + r) for r in range(nr)])
    eta = np.array([150.0, 400.0, 900.0])
    vt = np.stack([W.meridional_profile(t, 12.0 + 3.0 * r) for r in range(nr)])
    out = []
    for graphs in (0, 1):
        g = FluxTransport(t, p, nr, mode="independent")
        g.set_state(maps, layout="pt")
        g.set_diffusion(eta)
        g.set_flow_profiles(vt, None)
        g.set_flow_analytic(*W.dr_coefficients())
        # RKG2 super-step graphs (the host-controlled step; no planned step)
        g.set_tuning(fused=fused, fused_min_realizations=1, graphs=graphs, step_graph=0)
        states = []
        for i in range(6):
            g.flux_transport_step(1.0, strang=True, subcycles=2 if i < 4 else 0)
            states.append(g.get_state("tp"))
        if graphs:
            assert g.get_tuning("graph_builds") >= 1
            assert g.get_tuning("graph_hits") >= 1
        out.append(states)
        g.close()
    for a, b in zip(out[1], out[0]):
        assert_same(a, b, f"graph replay (fused={fused})")


def test_graph_replay_after_coef_table_growth(gpu_available):
    """A super-step whose stage table outgrows the device coefficient table
    (8 realizations x 146 stages > 1024 entries) reallocates it; graphs
    captured before hold the old table's address, so they must be dropped:
    repeating the first step's stage counts afterwards captures anew and
    matches direct launches bit for bit."""
    t, p = grid(91, 181)
    nr = 8
    maps = np.stack([W.bipole_map(t, p, 20, 400 + r) for r in range(nr)])
    eta = np.linspace(150.0, 900.0, nr)
    dts = [1.0, 1.0, 200.0, 1.0, 1.0]
    out, builds = [], None
    for graphs in (0, 1):
        g = FluxTransport(t, p, nr, mode="independent")
        g.set_state(maps, layout="pt")
        g.set_diffusion(eta)
        g.set_tuning(fused_min_realizations=1, graphs=graphs, step_graph=0)
        states = []
        for dt in dts:
            g.diffusion_step(dt, subcycles=1)
            states.append(g.get_state("tp"))
        s, _ = g.diffusion_stats()
        if graphs:
            builds = g.get_tuning("graph_builds")
            hits = g.get_tuning("graph_hits")
        out.append(states)
        g.close()
    assert s.max() == 11
    # captures: dt=1 (replayed once), dt=200, dt=1 again after the table grew
    assert builds == 3 and hits == 2, (builds, hits)
    for i, (a, b) in enumerate(zip(out[1], out[0])):
        assert_same(a, b, f"graph replay after table growth, step {i}")


def test_driver_writes_map_files(gpu_available, tmp_path):
    """The driver's map output (write_map / output cadence): initial, indexed
    and final HDF5 maps in the reference's layout; the final file holds the
    final state (float32 by default, float64 on request)."""
    from hipft_amd import h5io
    cfg = dict(res_nt=31, res_np=61, n_realizations=2, validation_run=1, advance_diffusion=True,
               diffusion_coef_constants=[500.0, 800.0], time_end=3.0, dt_max=1.0,
               output_map_time_cadence=1.0, output_map_directory=str(tmp_path),
               output_single_precision=False)
    sol, num, R = driver.run(cfg)
    names = sorted(os.listdir(tmp_path))
    assert "hipft_brmap_initial.h5" in names and "hipft_brmap_final.h5" in names
    assert "hipft_brmap_idx000001.h5" in names and "hipft_brmap_idx000004.h5" in names
    d, sc = h5io.read_map(str(tmp_path / "hipft_brmap_final.h5"))
    assert d.shape == (2, 31, 61) and np.array_equal(sc[2], [1.0, 2.0])
    assert np.array_equal(d, R.model.get_state("pt"))
    d0, _ = h5io.read_map(str(tmp_path / "hipft_brmap_initial.h5"))
    assert not np.array_equal(d0, d)


def test_driver_main_writes_reference_outputs(gpu_available, tmp_path):
    """driver.main, the counterpart of the reference executable, on the
    diffuse_soccer namelist cut to 48 h: the history files hold the
    reference's first rows, hipft_realization_parameters.out the four
    diffusivities, and hipft_timing.out timers that add up (the diffusion
    phase inside the flux transport, no advection or source time)."""
    ref = golden("diffuse_soccer")
    text = str(ref["hipft_in"]).replace("time_end = 672.0", "time_end = 48.0")
    assert "time_end = 48.0" in text
    (tmp_path / "hipft.in").write_text(text)
    out = tmp_path / "out"
    out.mkdir()
    sol, num, tm = driver.main(str(tmp_path / "hipft.in"), str(out))
    n = sol.shape[1]
    assert num[0, -1, 1] == 48.0
    assert np.array_equal(num[..., 0], ref["num"][:, :n, 0])
    err, verr, _ = history_error(sol, ref["sol"][:, :n])
    assert err <= 5e-12 and verr <= 1e-6
    for r in range(4):
        for kind, arr in (("sol", sol), ("num", num)):
            rows = np.loadtxt(out / f"hipft_history_{kind}_r{r + 1:06d}.out", skiprows=1)
            np.testing.assert_allclose(rows, arr[r], rtol=1e-15, atol=0)
    params = np.loadtxt(out / "hipft_realization_parameters.out", skiprows=1)
    assert np.array_equal(params[:, 11], [100.0, 200.0, 400.0, 800.0])
    assert 0.0 < tm["diffusion"] <= tm["flux_transport"] + 1e-3
    assert tm["advection"] == 0.0 and tm["source"] == 0.0
    assert tm["total"] >= tm["setup"] + tm["flux_transport"]
    summary = {}
    for line in (out / "hipft_timing.out").read_text().split("\n"):
        if line.startswith(" --> Flux transport:    ") or line.startswith(" Wall clock time:    "):
            summary[line[:25].strip()] = [float(x) for x in line[25:].split()]
    assert summary["Wall clock time:"][0] == round(tm["total"], 3)
    assert summary["--> Flux transport:"][:3] == [round(tm["flux_transport"], 3)] * 3


def test_driver_flows_from_files(gpu_available, tmp_path):
    """The driver with use_flow_from_files: a flow list of three phi-dependent
    flow pairs (HDF5, m/s) switched at 2.4 h and 6 h and wrapped at 9.6 h
    (time steps cut to those times, update_timestep 4410-4420), plus analytic
    differential rotation, Strang splitting and diffusion.  Replayed on the
    oracle step by step (same dt sequence, same active flow): the CFL limit of
    every step and the final state agree bit for bit."""
    from hipft_amd import h5io, validation as V
    from hipft_amd.inputs import FlowFiles
    s1, s2 = V.uniform_scales(31, 61)
    ntm, npm1 = len(s2), len(s1)
    nt, npm = ntm + 1, npm1 + 1
    rng = np.random.default_rng(5)
    th = np.linspace(0.0, np.pi, nt)
    names = []
    for i in range(3):
        a, ph = rng.uniform(10.0, 20.0), rng.uniform(0, 2 * np.pi)
        vt = -a * np.sin(2 * th)[:, None] * (1.0 + 0.3 * np.cos(s1[None, :] + ph))
        pp = np.linspace(0, 2 * np.pi, npm)
        vp = 30.0 * np.sin(s2)[:, None] * np.sin(2 * pp[None, :] + ph)
        vt[:, 0] = vt[:, -1]
        vp[:, 0], vp[:, -1] = vp[:, -2], vp[:, 1]
        h5io.write_map(str(tmp_path / f"vt{i}.h5"), vt, [s1, th], single=False)
        h5io.write_map(str(tmp_path / f"vp{i}.h5"), vp, [pp, s2], single=False)
        names.append((f"vt{i}.h5", f"vp{i}.h5"))
    with open(tmp_path / "flows.csv", "w") as fh:
        fh.write("DAY        VT          VP\n")
        for d, (a, b) in zip([0.0, 0.1, 0.25], names):
            fh.write(f"{d:11.5f} {a:11s} {b:11s}\n")
    nr = 2
    cfg = dict(res_nt=31, res_np=61, n_realizations=nr, validation_run=1, advance_flow=True,
               flow_dr_model=1, use_flow_from_files=True, flow_list_filename="flows.csv",
               flow_root_dir=".", base_dir=str(tmp_path), advance_diffusion=True,
               diffusion_coef_constants=[300.0, 600.0], strang_splitting=True, time_end=11.0,
               dt_max=1.5)
    sol, num, R = driver.run(cfg)
    o = Oracle(R.t, R.p, nr)
    o.ingest_pt(np.broadcast_to(V.soccer_ball(s1, s2), (nr, ntm, npm1)))
    o.load_diffusion([300.0, 600.0])
    ff = FlowFiles(str(tmp_path / "flows.csv"), str(tmp_path), 0.0, ntm, npm1)
    time = 0.0
    used = set()
    for row in num[0][1:]:
        if ff.update(time):
            vt, vp = ff.arrays()
        used.add(ff.idx)
        o.flow_reset()
        o.flow_set(np.broadcast_to(vt, (nr, npm, nt)), np.broadcast_to(vp, (nr, npm, ntm)))
        o.flow_add_dr([46.0] * nr, [-262.0] * nr, [-379.0] * nr)
        o.set_lf_alpha()
        assert o.flow_dtmax(1) == row[3], f"CFL limit at t={time}"
        dt = row[2]
        o.advection_step(0.5 * dt)
        o.diffusion_step(dt, 0)
        o.advection_step(0.5 * dt)
        time = time + dt
    assert used == {0, 1, 2}
    times = list(num[0][:, 1])
    for tc in (2.4, 6.0, 9.6):
        assert any(abs(x - tc) < 1e-12 for x in times), (tc, times)
    assert_same(R.model.get_state("tp"), o.get_f(), "driver with flows from files")


@pytest.mark.parametrize("model", [1, 2])
def test_rfe_frames_match_oracle(gpu_available, model):
    """Random flux emergence frames from the same random map: scaling, pole
    averages (sequential sum as generate_rfe) and flux balance.  The balance
    ratio comes from block partial sums on the GPU and the sequential get_flux
    order in the oracle, so frames agree to 1e-14."""
    nr = 3
    o, g = pair(45, 91, nr=nr)
    rng = np.random.default_rng(21)
    uf = [150.0, 60.0, 300.0]
    for slot in (0, 1):
        rnd = (np.pi / np.sqrt(2 * np.pi)) * rng.standard_normal((o.npm - 2, o.ntm))
        g.rfe_generate(slot, rnd, model, uf)
        assert_same(g.get_rfe(slot), o.rfe_frame(rnd, model, uf), f"RFE frame model {model}",
                    1e-14)


def _source_setup(o, g, nr, seed):
    rng = np.random.default_rng(seed)
    src = 1e-3 * rng.standard_normal((o.npm, o.ntm))
    src[0], src[-1] = src[-2], src[1]
    g.set_source(src)
    for slot in (0, 1):
        rnd = (np.pi / np.sqrt(2 * np.pi)) * rng.standard_normal((o.npm - 2, o.ntm))
        g.rfe_generate(slot, rnd, 1, [100.0 + 50.0 * r for r in range(nr)])
    fr = [g.get_rfe(0), g.get_rfe(1)]
    return src, fr


def _blend(src, fr, first, alpha):
    """update_source's source term: ((0 + map) + (1 - alpha) rfe1) + alpha rfe2."""
    s = 0.0 + src[None]
    s = s + (1.0 - alpha) * fr[first]
    return s + alpha * fr[1 - first]


def test_source_step_bitexact(gpu_available):
    nr = 2
    o, g = pair(45, 91, nr=nr)
    src, fr = _source_setup(o, g, nr, 5)
    for first, alpha, dt in ((0, 0.25, 0.7), (1, 0.8, 1.3)):
        g.rfe_weights(first, alpha)
        g.source_step(dt)
        o.source_step(dt, _blend(src, fr, first, alpha))
        assert_same(g.get_state("tp"), o.get_f(), f"source step first={first}")
    g.rfe_weights(-1, 0.0)
    g.source_step(0.5)
    o.source_step(0.5, np.broadcast_to(0.0 + src[None], (nr,) + src.shape))
    assert_same(g.get_state("tp"), o.get_f(), "source step, map only")


def test_flux_transport_step_with_sources(gpu_available):
    """Strang splitting with sources: S(dt/2) A(dt/2) D(dt) A(dt/2) S(dt/2)
    (flux_transport_step, src/hipft.f90:4490-4499), bit for bit."""
    nr = 2
    o, g = pair(61, 121, nr=nr)
    _flows(o, g, nr)
    eta = [300.0, 700.0]
    o.load_diffusion(eta)
    g.set_diffusion(eta)
    src, fr = _source_setup(o, g, nr, 9)
    g.rfe_weights(1, 0.4)
    s = _blend(src, fr, 1, 0.4)
    dt = 1.0
    g.flux_transport_step(dt, strang=True, subcycles=0)
    o.source_step(0.5 * dt, s)
    o.advection_step(0.5 * dt)
    o.diffusion_step(dt, 0)
    o.advection_step(0.5 * dt)
    o.source_step(0.5 * dt, s)
    assert_same(g.get_state("tp"), o.get_f(), "Strang step with sources")


def test_driver_random_flux_emergence(gpu_available):
    """A run with random flux emergence (model 1): frames renewed every
    lifetime with time steps cut to at most half a lifetime (update_timestep
    4431-4442); the balanced sources keep the signed flux at round-off while
    the unsigned flux grows; the same seed gives the same run."""
    cfg = dict(res_nt=31, res_np=61, n_realizations=2, advance_source=True, source_rfe_model=1,
               source_rfe_total_unsigned_flux_per_hours=[200.0, 50.0], source_rfe_lifetime=0.3,
               source_rfe_seed=12345, advance_diffusion=True,
               diffusion_coef_constants=[300.0, 300.0], strang_splitting=True, time_end=1.5,
               dt_max=1.0)
    sol, num, R = driver.run(cfg)
    dts = num[0][1:, 2]
    assert dts.max() <= 0.15 + 1e-12 and abs(num[0][-1, 1] - 1.5) < 1e-12
    fp, fm = sol[:, -1, 5], sol[:, -1, 6]
    assert np.all(np.abs(fp + fm) <= 1e-9 * fp)
    assert fp[0] > 2.0 * fp[1] > 0.0  # four times the budget, more unsigned flux
    sol2, _, _ = driver.run(cfg)
    assert np.array_equal(sol, sol2)


def test_wide_grid_fallback_kernels(gpu_available):
    """More than 1024 unique phi columns: the band and fused kernels (one
    thread or thread pair per column) give way to the pointwise stage
    kernels; a Strang step still matches the oracle bit for bit."""
    o, g = pair(33, 1300, nr=2)
    assert g.npm - 2 > 1024
    _flows(o, g, 2)
    eta = [400.0, 900.0]
    o.load_diffusion(eta)
    g.set_diffusion(eta)
    g.set_tuning(fused=1, advection_band=1)
    g.flux_transport_step(1.0, strang=True, subcycles=0)
    o.advection_step(0.5)
    o.diffusion_step(1.0, 0)
    o.advection_step(0.5)
    assert_same(g.get_state("tp"), o.get_f(), "Strang step at 33x1300")


def test_fortran_driver_matches_python(gpu_available, tmp_path):
    """The Fortran time loop (hipft_amd/fortran/hft_run.f90, through the
    iso_c_binding module include/hft.f90) gives the same history rows and
    final state, bit for bit, as the same run driven from Python."""
    import subprocess
    from hipft_amd import build as B
    exe = B.FORTRAN_DRIVER
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (no Fortran compiler)")
    ntm, npm1, nr, nsteps = 41, 81, 3, 4
    t, p = grid(ntm, npm1)
    maps = np.stack([W.bipole_map(t, p, 12, 70 + r) for r in range(nr)])
    eta = np.array([200.0, 500.0, 900.0])
    dr = np.array([W.dr_coefficients(1.0 + 0.05 * r) for r in range(nr)])
    mf = np.array([[22.0, 11.0, -28.0]] * nr)
    for name, a in (("t", t), ("p", p), ("map", maps)):
        np.ascontiguousarray(a, dtype=np.float64).tofile(str(tmp_path / f"{name}.bin"))
    fl = lambda v: ", ".join(repr(float(x)) for x in v)
    ctl = f"""&hft_run
  ntm = {ntm}, npm1 = {npm1}, nr = {nr}, mode = 1, nsteps = {nsteps}, strang = 1,
  subcycles = 0, dt = 1.0, advect = .true.,
  eta = {fl(eta)},
  dr_p0 = {fl(dr[:, 0])}, dr_p2 = {fl(dr[:, 1])}, dr_p4 = {fl(dr[:, 2])},
  mf_p1 = {fl(mf[:, 0])}, mf_p3 = {fl(mf[:, 1])}, mf_p5 = {fl(mf[:, 2])},
  t_file = 't.bin', p_file = 'p.bin', map_file = 'map.bin',
  hist_file = 'hist.bin', final_file = 'final.bin'
/
"""
    (tmp_path / "run.nml").write_text(ctl)
    r = subprocess.run([exe, "run.nml"], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = np.fromfile(str(tmp_path / "hist.bin")).reshape(nsteps, nr, 16)
    final = np.fromfile(str(tmp_path / "final.bin")).reshape(nr, npm1 + 1, ntm)
    g = FluxTransport(t, p, nr, mode="independent")
    g.set_state(maps, layout="pt")
    g.set_flow_analytic(p0=dr[:, 0], p2=dr[:, 1], p4=dr[:, 2], p1=mf[:, 0], p3=mf[:, 1],
                        p5=mf[:, 2])
    g.set_diffusion(eta)
    for n in range(nsteps):
        g.flux_transport_step(1.0, strang=True, subcycles=0)
        h = g.analysis()
        assert_same(rows[n, :, 2:], h, f"history row {n + 1}")
    assert_same(final, g.get_state("tp"), "final state")
    g.close()


def test_weno_eps_matches_oracle(gpu_available):
    """weno_eps > 0 (load_weno 6895-6901): every smoothness weight uses the
    constant instead of the cell widths; operator and SSPRK(4,3) steps equal
    the oracle bit for bit, and eps = 0 restores the cell widths."""
    o, g = pair(61, 121, nr=2)
    _flows(o, g, 2)
    x = o.get_f()
    base = g.apply_advection_operator(x)
    for eps in (1.0e-3, 0.5):
        o.set_weno_eps(eps)
        g.set_weno_eps(eps)
        y = g.apply_advection_operator(x)
        assert_same(y, o.advection_operator(x), f"A x, weno_eps={eps}")
        assert not np.array_equal(y, base)
    for _ in range(2):
        o.advection_step(0.5)
        g.advection_step(0.5)
    assert_same(g.get_state("tp"), o.get_f(), "SSPRK(4,3) with weno_eps")
    g.set_weno_eps(0.0)
    assert_same(g.apply_advection_operator(x), base, "weno_eps = 0 restores cell widths")


def test_pole_flux_lat_limit_matches_oracle(gpu_available):
    """Polar flux and area columns with pole_flux_lat_limit = 45 degrees."""
    o, g = pair(91, 181, nr=2)
    o.set_pole_flux_lat_limit(45.0)
    g.set_pole_flux_lat_limit(45.0)
    ho, hg = o.analysis(), g.analysis()
    scale = np.maximum(np.abs(ho[:, 3:5]).max(), 1.0)
    assert np.abs(hg[:, 3:11] - ho[:, 3:11]).max() <= 1e-13 * scale
    g.set_pole_flux_lat_limit(30.0)
    assert np.abs(g.analysis()[:, 9] - ho[:, 9]).max() > 1e-3 * abs(ho[0, 9])


def test_balance_flux_matches_oracle(gpu_available):
    """balance_flux (8752-8801) of each realization: in place, and as a PT
    copy that leaves the state unchanged; the GPU sums the fluxes in blocks,
    the oracle in sequence, hence 1e-14."""
    o, g = pair(61, 121, nr=3)
    f0 = g.get_state("tp")
    want = np.stack([o.balance_flux(f0[r]) for r in range(3)])
    cp = g.balance_flux(copy=True)
    assert np.array_equal(g.get_state("tp"), f0)
    tp_of_copy = np.transpose(cp, (0, 2, 1))  # (nr, npm1, ntm): columns 1..npm1
    assert_same(tp_of_copy, want[:, :o.npm1], "balanced copy", 1e-14)
    g.balance_flux()
    assert_same(g.get_state("tp"), want, "balanced in place", 1e-14)
    for r in range(3):
        fp, fm = o.get_flux(g.get_state("tp")[r])
        assert abs(fp + fm) <= 1e-13 * (fp - fm)


def test_driver_initial_balance_and_mult(gpu_available):
    """initial_map_flux_balance: ingest, balance, then initial_map_mult_fac
    (load_initial_condition 2140-2173), against the oracle."""
    t, p = grid(61, 121, clustered=False)
    maps = np.stack([W.bipole_map(t, p, 8, 40 + r) + 3.0 * (r + 1) for r in range(2)])
    R = driver.Run(dict(initial_map=maps, initial_scales=[p, t], n_realizations=2,
                        initial_map_flux_balance=True, initial_map_mult_fac=1.5))
    o = Oracle(t, p, 2)
    o.ingest_pt(maps)
    want = 1.5 * np.stack([o.balance_flux(f) for f in o.get_f()])
    # flux sums blocked on the GPU, sequential in the oracle: the balance
    # ratio differs by a few ulps of the sums
    assert_same(R.model.get_state("tp"), want, "balanced initial map", 1e-13)
    R.model.close()


def test_driver_index_cadence_map_list_and_balanced_output(gpu_available, tmp_path):
    """driver.main with output_map_idx_cadence = 2 and
    output_map_flux_balance: indexed maps every second step in
    output_map_directory (the initial condition first), the initial and final
    maps in the run directory, hipft_output_map_list.out with the reference's
    header and rows (i10, 1pe25.15), written maps flux balanced while the
    state (whose net flux the histories show) is not; pole_flux_lat_limit
    reaches the analysis."""
    from hipft_amd import h5io
    t, p = grid(31, 61, clustered=False)
    m0 = W.bipole_map(t, p, 6, 7) + 4.0  # net positive flux
    h5io.write_map(str(tmp_path / "init.h5"), m0, [p, t], single=False)
    text = ("&hipft_input_parameters\n initial_map_filename = 'init.h5'\n n_realizations = 2\n"
            " advance_diffusion = .true.\n"
            " diffusion_coef_constants = 500.0, 900.0\n time_end = 5.0\n dt_max = 1.0\n"
            " output_map_idx_cadence = 2\n output_map_flux_balance = .true.\n"
            " output_single_precision = .false.\n pole_flux_lat_limit = 40.0\n/\n")
    (tmp_path / "hipft.in").write_text(text)
    out = tmp_path / "run"
    out.mkdir()
    sol, num, tm = driver.main(str(tmp_path / "hipft.in"), str(out))
    assert sorted(os.listdir(out / "output_maps")) == [
        f"hipft_brmap_idx{i:06d}.h5" for i in (1, 2, 3)]
    assert (out / "hipft_brmap_initial.h5").exists() and (out / "hipft_brmap_final.h5").exists()
    lines = (out / "hipft_output_map_list.out").read_text().split("\n")
    assert lines[0] == "       IDX                      TIME FILENAME"
    assert lines[1] == "         1" + "   " + "    0.000000000000000E+00" + "    hipft_brmap_idx000001.h5"
    assert lines[2] == "         2" + "   " + "    2.000000000000000E+00" + "    hipft_brmap_idx000002.h5"
    assert lines[3] == "         3" + "   " + "    4.000000000000000E+00" + "    hipft_brmap_idx000003.h5"
    d, sc = h5io.read_map(str(out / "hipft_brmap_final.h5"))
    o = Oracle(t, p, 2)
    o.ingest_pt(d)
    for r in range(2):
        fp, fm = o.get_flux(o.get_f()[r])
        assert abs(fp + fm) <= 1e-13 * (fp - fm)
        # the state itself keeps its net flux (FLUX_POSITIVE + FLUX_NEGATIVE)
        assert sol[r, -1, 5] + sol[r, -1, 6] > 0.1 * sol[r, -1, 5]
    assert (out / "hipft_run_parameters_used.out").exists()
    o.set_pole_flux_lat_limit(40.0)
    ha = o.analysis()
    assert np.allclose(sol[:, -1, 11:13], ha[:, 9:11], rtol=1e-13, atol=0)
    o.set_pole_flux_lat_limit(30.0)
    assert not np.allclose(sol[:, -1, 11], o.analysis()[:, 9], rtol=1e-3)


PERSIST_OFF = ("the persistent super-step did not complete on the MI355X (the 91x181 six-"
               "realization case ran past its time limit twice, also with atomic polling and a "
               "wall-clock bound on every wait); cause not found, path switched off")


@pytest.mark.skip(reason=PERSIST_OFF)
@pytest.mark.parametrize("T,rows,long_rows,nlong", [(4, 32, 32, 0), (4, 64, 16, 2), (3, 16, 16, 0),
                                                    (2, 24, 8, 1)])
def test_persistent_superstep_matches_launches(gpu_available, T, rows, long_rows, nlong):
    """One persistent launch per RKG2 super-step (tiles of realization x band
    x stage group taken from a device list, each waiting for the previous
    group of its band and the bands either side) gives the multi-launch
    path's states bit for bit: Strang steps with PTL sub-cycling, six
    realizations with different stage counts, uniform and per-realization
    band heights, 2-4 stages per group."""
    t, p = grid(91, 181)
    nr = 6
    maps = np.stack([W.bipole_map(t, p, 20, 500 + r) for r in range(nr)])
    eta = np.array([900.0, 150.0, 420.0, 260.0, 700.0, 330.0])
    vt = np.stack([W.meridional_profile(t, 11.0 + 2.0 * r) for r in range(nr)])
    out = []
    for persist in (0, 1):
        g = FluxTransport(t, p, nr, mode="independent")
        g.set_state(maps, layout="pt")
        g.set_diffusion(eta)
        g.set_flow_profiles(vt, None)
        g.set_flow_analytic(*W.dr_coefficients())
        g.set_tuning(fused_stages=T, persist=persist, persist_rows=rows,
                     persist_rows_long=long_rows, persist_long=nlong)
        states = []
        for _ in range(3):
            g.flux_transport_step(1.0, strang=True, subcycles=0)
            states.append(g.get_state("tp"))
        s, _ = g.diffusion_stats()
        out.append(states)
        g.close()
    assert len(set(s.tolist())) > 2
    for i, (a, b) in enumerate(zip(out[1], out[0])):
        assert_same(a, b, f"persistent super-step, step {i}")


@pytest.mark.skip(reason=PERSIST_OFF)
def test_persistent_superstep_full_size(gpu_available):
    """512x1024, eight realizations (stage counts from ~110 to ~350): the
    persistent super-step equals the multi-launch path bit for bit."""
    t, p = grid(512, 1024)
    nr = 8
    maps = np.stack([W.bipole_map(t, p, 40, 600 + r) for r in range(nr)])
    eta = np.exp(np.linspace(np.log(100.0), np.log(1000.0), nr))
    out = []
    for persist in (0, 1):
        g = FluxTransport(t, p, nr, mode="independent")
        g.set_state(maps, layout="pt")
        g.set_diffusion(eta)
        g.set_tuning(persist=persist, persist_rows=64, persist_rows_long=32, persist_long=2)
        g.diffusion_step(1.0, subcycles=0)
        out.append(g.get_state("tp"))
        g.close()
    assert_same(out[1], out[0], "persistent super-step at 512x1024")


@pytest.mark.parametrize("case,hours", [("diffuse_soccer", 48.0),
                                        ("advect_gaussians_phi_theta", 24.0),
                                        ("diffuse_advect_soccer", 24.0)])
def test_fortran_hipft_executable_matches_python_driver(gpu_available, tmp_path, case, hours):
    """The Fortran executable (hipft_amd/fortran/hft_hipft.f90: reads the
    testsuite's hipft.in, runs the time loop through include/hft.f90, writes
    the history files) against driver.main on the same namelist cut to a
    day or two: the same rows (cycle rows at verbose = 2 included), the same
    header and line layout, values to 1e-13 (the validation fields are
    evaluated by Fortran and numpy, whose cos may differ in the last bit);
    the soccer-ball case also against the reference's own rows."""
    import subprocess
    from hipft_amd import build as B
    exe = B.FORTRAN_HIPFT
    if not os.path.exists(exe):
        pytest.skip("Fortran executable not built (no Fortran compiler)")
    ref = golden(case)
    text = str(ref["hipft_in"]).replace("time_end = 672.0", f"time_end = {hours}")
    assert f"time_end = {hours}" in text
    (tmp_path / "hipft.in").write_text(text)
    dpy, dfo = tmp_path / "py", tmp_path / "fo"
    dpy.mkdir()
    dfo.mkdir()
    driver.main(str(tmp_path / "hipft.in"), str(dpy))
    r = subprocess.run([exe, str(tmp_path / "hipft.in")], cwd=str(dfo), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nr = int(ref["sol"].shape[0])
    for k in range(1, nr + 1):
        for kind in ("sol", "num"):
            name = f"hipft_history_{kind}_r{k:06d}.out"
            a = (dfo / name).read_text().split("\n")
            b = (dpy / name).read_text().split("\n")
            assert a[0] == b[0] and len(a) == len(b), (name, a[0], b[0], len(a), len(b))
            assert [len(x) for x in a] == [len(x) for x in b], name
            fa = np.loadtxt(dfo / name, skiprows=1)
            fb = np.loadtxt(dpy / name, skiprows=1)
            if kind == "num":
                assert np.array_equal(fa[:, [0, 6, 7]], fb[:, [0, 6, 7]]), name
                assert np.allclose(fa, fb, rtol=1e-14, atol=0), name
            else:
                err, verr, _ = history_error(fa[None], fb[None])
                assert err <= 1e-13 and verr <= 1e-9, (name, err, verr)
    if case == "diffuse_soccer":
        sol = np.stack([np.loadtxt(dfo / f"hipft_history_sol_r{k:06d}.out", skiprows=1)
                        for k in range(1, nr + 1)])
        n = sol.shape[1]
        err, verr, _ = history_error(sol, ref["sol"][:, :n])
        assert err <= 5e-12 and verr <= 1e-6
