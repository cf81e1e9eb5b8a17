"""HDF5 map files (hipft_amd/h5io.py): the reference's layout (Data +
dim1..dim3 dimension scales, phi fastest, float32 unless double precision),
round trips, and the namelist's initial_map_filename."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import golden
from hipft_amd import h5io, namelist


def test_round_trip_single_and_double(tmp_path):
    g = golden("diffuse_dipole")
    data, sp, st = g["map_pt"].astype(np.float64), g["scale_p"], g["scale_t"]
    for single in (True, False):
        path = str(tmp_path / f"m{int(single)}.h5")
        h5io.write_map(path, data, [sp, st], single=single)
        d2, sc = h5io.read_map(path)
        want = data.astype(np.float32).astype(np.float64) if single else data
        assert d2.shape == (512, 1024)
        assert np.array_equal(d2, want)
        assert np.array_equal(sc[0], sp.astype(np.float32).astype(np.float64) if single else sp)
        assert np.array_equal(sc[1], st.astype(np.float32).astype(np.float64) if single else st)


def test_three_d_realization_scale(tmp_path):
    rng = np.random.default_rng(0)
    d = rng.standard_normal((3, 7, 12))
    path = str(tmp_path / "m3.h5")
    h5io.write_map(path, d, [np.linspace(0, 6.28, 12), np.linspace(0, 3.14, 7),
                             np.array([1.0, 2.0, 3.0])], single=False)
    d2, sc = h5io.read_map(path)
    assert np.array_equal(d2, d)
    assert np.array_equal(sc[2], [1.0, 2.0, 3.0])


@pytest.mark.skipif(shutil.which("h5dump") is None and not os.path.exists("/opt/conda/bin/h5dump"),
                    reason="h5dump not available")
def test_dimension_scales_attached(tmp_path):
    """Data carries DIMENSION_LIST/DIMENSION_LABELS and dim1 is the fastest
    (phi) axis, as the reference's files are laid out."""
    path = str(tmp_path / "m.h5")
    h5io.write_map(path, np.zeros((4, 9)), [np.arange(9.0), np.arange(4.0)])
    exe = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = subprocess.run([exe, "-H", path], capture_output=True, text=True).stdout
    assert "DIMENSION_LIST" in out and "DIMENSION_LABELS" in out
    assert "SIMPLE { ( 4, 9 ) / ( 4, 9 ) }" in out
    assert 'DATASET "dim1"' in out and "SIMPLE { ( 9 ) / ( 9 ) }" in out
    assert "H5T_IEEE_F32LE" in out


def test_namelist_initial_map(tmp_path):
    d = np.arange(5 * 8, dtype=np.float64).reshape(5, 8)
    h5io.write_map(str(tmp_path / "br.h5"), d, [np.linspace(0, 6.283185307179586, 8),
                                                 np.linspace(0, 3.141592653589793, 5)],
                   single=False)
    text = """&hipft_input_parameters
      initial_map_filename = 'br.h5'
      time_end = 2.0
    /"""
    cfg = namelist.to_run_config(namelist.parse(text), base_dir=str(tmp_path))
    assert np.array_equal(cfg["initial_map"], d)
    assert len(cfg["initial_scales"]) == 2 and len(cfg["initial_scales"][0]) == 8


def test_missing_file_raises(tmp_path):
    with pytest.raises(h5io.H5Error):
        h5io.read_map(str(tmp_path / "nope.h5"))
