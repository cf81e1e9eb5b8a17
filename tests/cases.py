"""The reference testsuite cases (testsuite/<case>/input/hipft.in) as
namelist dicts, plus the history comparison rule used by the parity tests."""
import numpy as np

from conftest import golden

RSUN_CM2 = 4.84416e21


def case_cfg(case):
    base = dict(verbose=2, res_nt=180, res_np=360, time_end=672.0, dt_max=1.0)
    if case == "advect_gaussians_phi":
        base.update(validation_run=2, advance_flow=True, flow_vp_rigid_omega=-9999.0)
    elif case == "advect_gaussians_theta":
        base.update(validation_run=2, advance_flow=True, flow_vt_const=-9999.0)
    elif case == "advect_gaussians_phi_theta":
        base.update(validation_run=2, advance_flow=True, flow_vp_rigid_omega=-9999.0,
                    flow_vt_const=-9999.0)
    elif case == "diffuse_soccer":
        base.update(n_realizations=4, validation_run=1, advance_diffusion=True,
                    diffusion_coef_constants=[100.0, 200.0, 400.0, 800.0])
    elif case == "diffuse_advect_soccer":
        base.update(n_realizations=4, validation_run=1, advance_diffusion=True,
                    diffusion_coef_constants=[100.0, 200.0, 400.0, 800.0], advance_flow=True,
                    flow_vp_rigid_omega=1.8076624395655556)
    elif case == "diffuse_dipole":
        g = golden(case)
        base = dict(verbose=2, initial_map=g["map_pt"].astype(np.float64),
                    initial_scales=(g["scale_p"].astype(np.float64), g["scale_t"].astype(np.float64)),
                    diffusion_coef_constant=500.0, time_end=8760.0, advance_diffusion=True)
    elif case == "diffuse_advect_atten_map_1cr":
        g = golden(case)
        base = dict(verbose=2, initial_map=g["map_pt"].astype(np.float64),
                    initial_scales=(g["scale_p"].astype(np.float64), g["scale_t"].astype(np.float64)),
                    output_map_time_cadence=100.0, output_history_time_cadence=0.5,
                    n_realizations=5, time_end=672.0, dt_max=0.25, advance_flow=True,
                    flow_attenuate=True, flow_attenuate_value=300.0, flow_dr_model=1,
                    flow_mf_model=1, flow_dr_coef_p0_values=[41.4, 46.0, 50.6, 46.0, 46.0],
                    flow_dr_coef_p2_values=[-288.2, -262.0, -235.8, -262.0, -262.0],
                    flow_dr_coef_p4_values=[-416.0, -379.0, -341.1, -379.0, -379.0],
                    flow_mf_coef_p1_values=[22.0, 22.0, 22.0, 17.8, 24.2],
                    flow_mf_coef_p3_values=[11.0, 11.0, 11.0, 9.9, 12.1],
                    flow_mf_coef_p5_values=[-28.0, -28.0, -28.0, -30.8, -25.2],
                    advance_diffusion=True, diffusion_coef_constant=25.0)
    else:
        raise KeyError(case)
    return base


def history_error(sol, ref):
    """Per-column error of history rows sol vs ref ([nr][rows][16]), each
    column normalised by a physical scale: |Br| columns by max|Br|, flux and
    area columns by the unsigned flux, dipoles by the unsigned flux moment.
    Returns (max error over columns 2..14, max error of column 15 relative
    to its largest value, per-column errors)."""
    n = sol.shape[1]
    ref = ref[:, :n]
    brmax = np.maximum(np.abs(ref[..., 2]), np.abs(ref[..., 3])).max(axis=1, keepdims=True)
    uflux = (ref[..., 5] - ref[..., 6]).max(axis=1, keepdims=True)
    scale = np.empty(ref.shape[:2] + (16,))
    scale[..., 2:5] = brmax[..., None]
    scale[..., 5:11] = uflux[..., None]
    scale[..., 11:13] = np.abs(ref[..., 11:13]).max(axis=1, keepdims=True)
    scale[..., 13:15] = (uflux / RSUN_CM2 * 0.75 / np.pi)[..., None]
    err = np.abs(sol[..., 2:15] - ref[..., 2:15]) / scale[..., 2:15]
    # validation error: |f - f_exact| is small, so its rounding error is
    # amplified; compare against the column's largest value
    v = np.abs(sol[..., 15] - ref[..., 15]) / np.maximum(
        np.abs(ref[..., 15]).max(axis=1, keepdims=True), 1e-300)
    return float(err.max()), float(v.max()), err.max(axis=(0, 1))
