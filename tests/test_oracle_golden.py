"""Pin the CPU oracle to the reference's own testsuite histories
(testsuite/<case>/reference/hipft_history_{sol,num}_r*.out, committed as
tests/golden/<case>.npz by tests/golden/make_fixtures.py).

The oracle restates the reference's time loop (update_timestep, Strang/unsplit
splitting, PTL sub-cycles, RKG2, SSPRK(4,3)+WENO3, attenuated flows,
analysis) and must reproduce each case's history rows.  The integer columns
of the num history (RKG2 stage count, cycles) must match exactly.
"""
import numpy as np
import pytest

from cases import case_cfg, history_error
from conftest import golden
from oracle.oracle import run_case

PREFIX = {
    "advect_gaussians_phi": 40,
    "advect_gaussians_theta": 40,
    "advect_gaussians_phi_theta": 40,
    "diffuse_soccer": 30,
    "diffuse_advect_soccer": 20,
    "diffuse_advect_atten_map_1cr": 30,
}


@pytest.mark.parametrize("case", sorted(PREFIX))
def test_oracle_history_prefix(case):
    sol, num, _ = run_case(case_cfg(case), max_steps=PREFIX[case])
    ref = golden(case)
    n = sol.shape[1]
    assert n > 10
    err, verr, cols = history_error(sol, ref["sol"])
    assert err <= 2e-13, (case, cols)
    assert verr <= 1e-7
    rn = ref["num"][:, :n]
    assert np.array_equal(num[..., 0], rn[..., 0])
    assert np.array_equal(num[..., 6:8], rn[..., 6:8])
    assert np.allclose(num[..., 1:6], rn[..., 1:6], rtol=1e-14, atol=0)


def test_oracle_full_rotation_advect_gaussians_phi():
    """One full rigid rotation: all 672 steps against the reference rows."""
    case = "advect_gaussians_phi"
    sol, num, _ = run_case(case_cfg(case))
    ref = golden(case)
    assert sol.shape == ref["sol"].shape
    err, verr, cols = history_error(sol, ref["sol"])
    assert err <= 1e-12, cols
    assert np.array_equal(num[..., 6:8], ref["num"][..., 6:8])


def test_case_configs_match_reference_namelists():
    """The case dicts used by the tests say what the reference's hipft.in files say."""
    from hipft_amd import namelist
    keys = ["res_nt", "res_np", "n_realizations", "validation_run", "time_end", "dt_max",
            "advance_flow", "advance_diffusion", "flow_vp_rigid_omega", "flow_vt_const",
            "diffusion_coef_constant", "diffusion_coef_constants", "flow_attenuate",
            "flow_attenuate_value", "flow_dr_model", "flow_mf_model", "flow_dr_coef_p0_values",
            "flow_dr_coef_p2_values", "flow_dr_coef_p4_values", "flow_mf_coef_p1_values",
            "flow_mf_coef_p3_values", "flow_mf_coef_p5_values", "output_history_time_cadence",
            "output_map_time_cadence", "verbose"]
    for case in list(PREFIX) + ["diffuse_dipole"]:
        nml = namelist.parse(str(golden(case)["hipft_in"]))
        cfg = case_cfg(case)
        for k in keys:
            if k in nml:
                a, b = nml[k], cfg.get(k)
                if isinstance(a, list):
                    assert list(map(float, a)) == list(map(float, b)), (case, k)
                else:
                    assert a == b, (case, k, a, b)
