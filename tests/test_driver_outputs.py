"""The driver's text outputs against the reference testsuite's own files.

* hipft_timing.out (write_timing, src/hipft.f90:1414-1552): the file the
  diffuse_soccer case holds is rewritten from timer values that print as its
  numbers in both of its formats (f20.2 and f12.3), byte for byte.
* hipft_realization_parameters.out (set_realization_parameters,
  src/hipft.f90:2346-2383): the parameters of the diffuse_soccer (four
  diffusivities) and diffuse_advect_atten_map_1cr (five flow coefficient sets)
  namelists, written by the driver, read back and compared with the files the
  cases hold.  Those files come from an earlier reference release (12 columns
  of width 29); the current writer has 13 columns of width 41 (the RFE flux
  column added), so the comparison is by column name and value.
"""
import numpy as np

import os

from hipft_amd import driver, namelist
from conftest import GOLDEN, golden


def golden_path(name):
    return os.path.join(GOLDEN, name)


def _read_table(text):
    lines = [ln for ln in text.split("\n") if ln.strip()]
    names = lines[0].split()
    rows = np.array([[float(x) for x in ln.split()] for ln in lines[1:]])
    return names, rows


def test_timing_file_matches_reference_layout(tmp_path):
    ref = str(golden("diffuse_soccer")["timing"])
    tm = dict(total=2.403, setup=0.031, update=0.001, data_assimilation=0.0,
              flux_transport=1.838, advection=0.0, diffusion=1.838, source=0.0,
              analysis=0.830, io=0.0249, mpi_overhead=0.094)
    path = driver.write_timing(tm, str(tmp_path))
    assert open(path).read() == ref


def _cfg(case):
    nml = namelist.parse(str(golden(case)["hipft_in"]))
    nml.pop("initial_map_filename", None)  # the parameters do not need the map
    return namelist.to_run_config(nml, ".")


def test_realization_parameters_match_reference(tmp_path):
    for case, nr in (("diffuse_soccer", 4), ("diffuse_advect_atten_map_1cr", 5)):
        cfg = _cfg(case)
        assert int(cfg["n_realizations"]) == nr
        path = driver.write_realization_parameters(cfg, str(tmp_path))
        text = open(path).read()
        names, rows = _read_table(text)
        ref_names, ref_rows = _read_table(str(golden(case)["realization_parameters"]))
        assert names[:len(ref_names)] == ref_names
        assert names[-1] == "source_rfe_total_unsigned_flux_per_hours"
        assert rows.shape == (nr, 13)
        np.testing.assert_array_equal(rows[:, :ref_rows.shape[1]], ref_rows)
        assert np.all(rows[:, 0] == np.arange(1, nr + 1))
        # (100(A41)) and (I41,a,100(1pe40.16e3,a))
        lines = text.split("\n")
        assert len(lines[0]) == 41 * 13
        assert all(len(ln) == 42 + 41 * 12 for ln in lines[1:nr + 1])
        assert lines[1][42:82] == "0.0000000000000000E+000".rjust(40)


def test_realization_parameters_single_realization(tmp_path):
    assert driver.write_realization_parameters({"n_realizations": 1}, str(tmp_path)) is None
    assert not (tmp_path / "hipft_realization_parameters.out").exists()


def test_fortran_e_format():
    assert driver._fortran_e(-262.0, 40, 16) == "-2.6200000000000000E+002".rjust(40)
    assert driver._fortran_e(1.0e-300, 23, 15) == "1.000000000000000E-300".rjust(23)
    assert driver._fortran_e(0.1, 40, 16) == "1.0000000000000001E-001".rjust(40)


def test_run_parameters_used_matches_reference(tmp_path):
    """hipft_run_parameters_used.out against the files two testsuite cases
    hold: every item the reference wrote has the same value here.  Those
    files come from an earlier release that wrote the per-realization arrays
    before filling them (sentinels -1 and -1e6); the current source fills
    unset entries with the scalar first (process_input_parameters,
    src/hipft.f90:8542-8579, before the write at 8450-8454), so array entries
    are compared through that rule."""
    from hipft_amd.namelist import REFERENCE_KEYS
    fx = np.load(golden_path("run_parameters_used.npz"))
    for case in ("diffuse_soccer", "diffuse_advect_atten_map_1cr"):
        nml = namelist.parse(str(golden(case)["hipft_in"]))
        path = driver.write_run_parameters(nml, str(tmp_path))
        ours = namelist.parse(open(path).read())
        assert list(ours) == list(REFERENCE_KEYS)
        ref = namelist.parse(str(fx[case]))
        for k, rv in ref.items():
            ov = ours[k]
            if k in driver.ARRAY_DEFAULTS:
                low = -100000.0 if k.startswith(("flow_dr", "flow_mf")) else 0.0
                scalar = ref[driver.ARRAY_DEFAULTS[k]]
                want = [v if v >= low else scalar for v in rv]
                assert len(ov) == len(rv) == driver.MAX_REALIZATIONS, k
                np.testing.assert_array_equal(ov, want, err_msg=k)
            elif isinstance(rv, float):
                assert ov == rv, (k, ov, rv)
            else:
                assert ov == rv, (k, ov, rv)


def test_namelist_refuses_unknown_and_unsupported_items():
    """An item outside &hipft_input_parameters is an error (as for a Fortran
    namelist read); reference options the library does not implement are
    refused instead of silently running another method."""
    import pytest
    base = "&hipft_input_parameters\n res_nt=31\n res_np=61\n time_end=2.0\n"
    with pytest.raises(ValueError, match="unknown namelist item"):
        namelist.to_run_config(namelist.parse(base + " no_such_item = 3\n/\n"))
    for item, msg in (("diffusion_num_method = 1", "diffusion_num_method"),
                      ("flow_num_method = 3", "flow_num_method"),
                      ("upwind = 0.5", "upwind"), ("flow_rigid_omega = 0.2", "flow_rigid_omega"),
                      ("diffusion_coef_filename = 'eta.h5'", "diffusion_coef_filename"),
                      ("diffusion_coef_grid = .true.", "diffusion_coef_grid"),
                      ("output_flows = .true.", "output_flows"),
                      ("output_map_time_cadence = 1.0\n output_map_idx_cadence = 2",
                       "cannot both")):
        cfg = namelist.to_run_config(namelist.parse(base + f" {item}\n/\n"))
        with pytest.raises(ValueError, match=msg):
            driver.check_config(dict(driver.DEFAULTS, **cfg))
    # supported values pass
    cfg = namelist.to_run_config(namelist.parse(
        base + " diffusion_num_method = 3\n flow_num_method = 4\n upwind = 1.0\n/\n"))
    driver.check_config(dict(driver.DEFAULTS, **cfg))


def test_per_realization_arrays_follow_the_sentinel_rule():
    """diffusion_coef_constants(2) = 300 sets one element; entries left at
    the sentinel (< 0; < -1e5 for flow coefficients) take the scalar."""
    text = ("&hipft_input_parameters\n n_realizations = 4\n diffusion_coef_constant = 50.\n"
            " diffusion_coef_constants(2) = 300.\n flow_dr_coef_p0_values = 10., -1000000., 30.\n"
            " flow_dr_coef_p0_value = 7.\n/\n")
    cfg = namelist.to_run_config(namelist.parse(text))
    np.testing.assert_array_equal(driver.per_realization(cfg, "diffusion_coef_constant", 4),
                                  [50.0, 300.0, 50.0, 50.0])
    np.testing.assert_array_equal(driver.per_realization(cfg, "flow_dr_coef_p0_value", 4),
                                  [10.0, 7.0, 30.0, 7.0])
