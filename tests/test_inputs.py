"""File lists of a run (hipft_amd/inputs.py): the flow list (flows from
files) and the assimilation map list.  Times and indices are checked against
values worked out by hand from load_flow_from_files (src/hipft.f90:3675-3860),
add_flow_from_files (3525-3673), load_data_assimilation (3949-4081) and
apply_data_assimilation (4283-4296)."""
import numpy as np
import pytest

from hipft_amd import h5io
from hipft_amd.inputs import (M_S_TO_RS_HR, AssimilationList, FlowFiles, read_assimilation_list,
                              read_flow_list)


def write_flow_list(path, jds, names):
    # '(F11.5,1X,A11,1X,A11)': the day numbers must fit 11 characters
    with open(path, "w") as fh:
        fh.write("JD         VT_FILE     VP_FILE\n")
        for jd, (a, b) in zip(jds, names):
            fh.write(f"{jd:11.5f} {a:11s} {b:11s}\n")


def flow_setup(tmp_path, jds):
    names = [(f"vt{i:03d}.h5", f"vp{i:03d}.h5") for i in range(len(jds))]
    write_flow_list(tmp_path / "flows.csv", jds, names)
    return str(tmp_path / "flows.csv")


def test_flow_list_parse(tmp_path):
    path = flow_setup(tmp_path, [100.0, 100.25, 100.5, 101.0])
    times, ft, fp = read_flow_list(path)
    assert times == [0.0, 6.0, 12.0, 24.0]
    assert ft[2] == "vt002.h5" and fp[3] == "vp003.h5"


@pytest.mark.parametrize("time_start,idx,current,nxt", [
    (0.0, 0, 0.0, 6.0),      # start of the list
    (13.0, 2, 12.0, 24.0),   # nearest time below the start
    (11.0, 1, 6.0, 12.0),    # nearest time above the start: the one before it
    (30.0, 1, 30.0, 36.0),   # the list repeats with period 24 h
    (24.0, 0, 24.0, 30.0),   # MODULO(24, 24) = 0
])
def test_flow_list_start(tmp_path, time_start, idx, current, nxt):
    path = flow_setup(tmp_path, [100.0, 100.25, 100.5, 101.0])
    ff = FlowFiles(path, str(tmp_path), time_start, 5, 9)
    assert (ff.idx, ff.time_current, ff.time_next) == (idx, current, nxt)


def test_flow_list_advance_and_wrap(tmp_path):
    path = flow_setup(tmp_path, [100.0, 100.25, 100.5, 101.0])
    ff = FlowFiles(path, str(tmp_path), 0.0, 5, 9)
    seq = []
    for t in (0.0, 3.0, 6.0, 9.0, 12.0, 24.0, 30.0, 36.0, 41.0, 42.0):
        changed = ff.update(t)
        ff.loaded = ff.idx
        seq.append((t, changed, ff.idx, ff.time_next))
    # the last entry lasts as long as the spacing of the last two (12 h),
    # then the list starts again
    assert seq == [(0.0, True, 0, 6.0), (3.0, False, 0, 6.0), (6.0, True, 1, 12.0),
                   (9.0, False, 1, 12.0), (12.0, True, 2, 24.0), (24.0, True, 3, 36.0),
                   (30.0, False, 3, 36.0), (36.0, True, 0, 42.0), (41.0, False, 0, 42.0),
                   (42.0, True, 1, 48.0)]


def test_flow_files_arrays(tmp_path):
    ntm, npm1 = 6, 11
    nt, npm = ntm + 1, npm1 + 1
    rng = np.random.default_rng(3)
    t_ms = rng.standard_normal((nt, npm1)) * 10.0
    p_ms = rng.standard_normal((ntm, npm)) * 10.0
    sp, st = np.linspace(0, 2 * np.pi, npm1), np.linspace(0, np.pi, nt)
    h5io.write_map(str(tmp_path / "vt000.h5"), t_ms, [sp, st], single=False)
    h5io.write_map(str(tmp_path / "vp000.h5"), p_ms, [np.linspace(0, 2 * np.pi, npm),
                                                    np.linspace(0, np.pi, ntm)], single=False)
    h5io.write_map(str(tmp_path / "vt001.h5"), t_ms, [sp, st], single=False)
    h5io.write_map(str(tmp_path / "vp001.h5"), p_ms, [np.linspace(0, 2 * np.pi, npm),
                                                    np.linspace(0, np.pi, ntm)], single=False)
    path = flow_setup(tmp_path, [100.0, 101.0])
    ff = FlowFiles(path, str(tmp_path), 0.0, ntm, npm1)
    vt, vp = ff.arrays()
    assert vt.shape == (1, npm, nt) and vp.shape == (1, npm, ntm)
    # vt(i,j) = m_s_to_rs_hr * file(j,i) (3596-3598); seam columns periodic
    for k in range(1, npm - 1):
        assert np.array_equal(vt[0, k], M_S_TO_RS_HR * t_ms[:, k])
        assert np.array_equal(vp[0, k], M_S_TO_RS_HR * p_ms[:, k])
    assert np.array_equal(vt[0, 0], vt[0, npm - 2]) and np.array_equal(vt[0, npm - 1], vt[0, 1])
    assert np.array_equal(vp[0, 0], vp[0, npm - 2]) and np.array_equal(vp[0, npm - 1], vp[0, 1])
    with pytest.raises(ValueError):
        FlowFiles(path, str(tmp_path), 0.0, ntm + 1, npm1).arrays()


def write_assim_list(path, jds, files):
    with open(path, "w") as fh:
        fh.write("obs_datetime_utc    target_datetime_utc obs_jd        path\n")
        for jd, f in zip(jds, files):
            fh.write(f"2022-01-01T00:00:00 2022-01-01T00:12:00 {jd:13.5f} {f}\n")


def test_assimilation_list_parse(tmp_path):
    write_assim_list(tmp_path / "maps.csv", [2459580.5, 2459581.0], ["a/m1.h5", "a/m2.h5"])
    req, act, jd, files = read_assimilation_list(str(tmp_path / "maps.csv"))
    assert jd == [2459580.5, 2459581.0] and files == ["a/m1.h5", "a/m2.h5"]
    assert req[0] == "2022-01-01T00:00:00" and act[1] == "2022-01-01T00:12:00"


@pytest.mark.parametrize("time_start,first", [
    (0.0, 0), (-3.0, 0), (5.0, 1), (7.0, 1), (12.0, 1), (13.0, 2), (24.0, 2), (25.0, None)])
def test_assimilation_list_start(tmp_path, time_start, first):
    write_assim_list(tmp_path / "maps.csv", [2459580.5, 2459581.0, 2459581.5],
                     ["m1.h5", "m2.h5", "m3.h5"])
    al = AssimilationList(str(tmp_path / "maps.csv"), str(tmp_path), time_start)
    assert al.first == first
    assert al.times == ([] if first is None else [0.0, 12.0, 24.0][first:])


def test_assimilation_list_maps(tmp_path):
    d = np.random.default_rng(1).standard_normal((3, 5, 9))
    h5io.write_map(str(tmp_path / "m2.h5"), d, [np.arange(9.0), np.arange(5.0),
                                                np.arange(1.0, 4.0)], single=False)
    write_assim_list(tmp_path / "maps.csv", [2459580.5, 2459581.0], ["m1.h5", "m2.h5"])
    al = AssimilationList(str(tmp_path / "maps.csv"), str(tmp_path), 12.0)
    assert np.array_equal(al.map(0), d)
