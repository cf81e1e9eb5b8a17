"""Time-dependent inputs of a run read from file lists: flows from files and
the data-assimilation map list.

Flow list (load_flow_from_files, src/hipft.f90:3675-3860; add_flow_from_files
3525-3673): a header row, then fixed-width rows '(F11.5,1X,A11,1X,A11)' of
(Julian day, vt file, vp file).  Times are 24*JD hours shifted to start at 0;
the list repeats with period = time span of the list, the step after the
last entry being the last entry's spacing.  vt files hold (npm1, nt) and vp
files (np, ntm) values in m/s (Fortran order, phi fastest).

Assimilation list (load_data_assimilation, src/hipft.f90:3949-4081;
apply_data_assimilation 4148-4306): a header row, then rows
'(A19,1X,A19,1X,F13.5,1X,A)' of (requested time, actual time, actual Julian
day, map file).  Map k is assimilated at 24*JD_k - 24*JD_1 hours.  Map files
hold (npm1, ntm, 3) slices: data, weight, mu.
"""
import math
import os

import numpy as np

M_S_TO_RS_HR = 5.172413793103448e-06  # src/hipft.f90:128


def _fortran_modulo(a, p):
    """Fortran MODULO for reals: a - floor(a/p)*p (sign of p)."""
    return a - math.floor(a / p) * p


def read_flow_list(path):
    """(times_hr starting at 0, vt file names, vp file names)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    rows = [ln for ln in lines[1:] if ln.strip()]
    jd, ft, fp = [], [], []
    for ln in rows:
        ln = ln.ljust(35)
        jd.append(float(ln[0:11]))
        ft.append(ln[12:23].strip())
        fp.append(ln[24:35].strip())
    times = [24.0 * x for x in jd]
    t0 = times[0]
    return [x - t0 for x in times], ft, fp


class FlowFiles:
    """Which flow file pair is active at a time, and its arrays in the
    storage hft_set_flow_arrays takes (code units)."""

    def __init__(self, list_path, root_dir, time_start, ntm, npm1):
        self.times, self.ft, self.fp = read_flow_list(list_path)
        self.root = root_dir
        self.ntm, self.npm1 = ntm, npm1
        n = len(self.times)
        if n < 2:
            raise ValueError(f"flow list {list_path}: at least two flows are needed")
        span = self.times[-1] - self.times[0]
        self.dt_last = self.times[-1] - self.times[-2]
        tsf = _fortran_modulo(time_start, span)
        d = [abs(x - tsf) for x in self.times]
        idx = d.index(min(d))  # first minimum (MINLOC)
        if self.times[idx] > tsf:
            idx -= 1
        self.idx = idx  # 0-based
        self.time_current = time_start - tsf + self.times[idx]
        self.time_next = self.time_current + self._delta()
        self.loaded = None

    def _delta(self):
        i = self.idx
        return self.dt_last if i + 1 >= len(self.times) else self.times[i + 1] - self.times[i]

    def update(self, time):
        """Advance the list to time (add_flow_from_files); True when the
        active pair changed or nothing is loaded yet."""
        changed = self.loaded is None
        if time >= self.time_next:
            self.idx += 1
            if self.idx >= len(self.times):
                self.idx = 0
            self.time_current = self.time_next
            self.time_next = self.time_current + self._delta()
            changed = True
        return changed

    def arrays(self):
        """(vt [1][npm][nt], vp [1][npm][ntm]) of the active pair, code units,
        the periodic seam columns set as set_periodic_bc_3d does."""
        from . import h5io
        ntm, npm1 = self.ntm, self.npm1
        npm, nt = npm1 + 1, ntm + 1
        dt_, _ = h5io.read_map(os.path.join(self.root, self.ft[self.idx]))
        dp_, _ = h5io.read_map(os.path.join(self.root, self.fp[self.idx]))
        if dt_.shape != (nt, npm1) or dp_.shape != (ntm, npm):
            raise ValueError(f"flow files {self.ft[self.idx]}, {self.fp[self.idx]}: shapes "
                             f"{dt_.shape}, {dp_.shape}; expected ({nt}, {npm1}), ({ntm}, {npm})")
        vt = np.empty((1, npm, nt))
        vt[0, :npm1] = M_S_TO_RS_HR * dt_.T
        vt[0, 0] = vt[0, npm - 2]
        vt[0, npm - 1] = vt[0, 1]
        vp = np.empty((1, npm, ntm))
        vp[0] = M_S_TO_RS_HR * dp_.T
        vp[0, 0] = vp[0, npm - 2]
        vp[0, npm - 1] = vp[0, 1]
        self.loaded = self.idx
        return vt, vp


def read_assimilation_list(path):
    """(requested time strings, actual time strings, Julian days, files)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    rows = [ln for ln in lines[1:] if ln.strip()]
    req, act, jd, files = [], [], [], []
    for ln in rows:
        req.append(ln[0:19])
        act.append(ln[20:39])
        jd.append(float(ln[40:53]))
        files.append(ln[54:].strip())
    return req, act, jd, files


class AssimilationList:
    """Times (hours from the first map) and map reader of an assimilation
    list, starting at the map load_data_assimilation picks for time_start."""

    def __init__(self, list_path, root_dir, time_start):
        _, _, jd, self.files = read_assimilation_list(list_path)
        self.root = root_dir
        hr = [24.0 * x for x in jd]
        t0, tf = hr[0], hr[-1]
        self.all_times = [x - t0 for x in hr]
        d = [abs((x - t0) - time_start) for x in hr]
        i = d.index(min(d))
        if time_start <= 0.0:
            first = 0
        elif time_start > tf - t0:
            first = None  # past the data: no assimilation
        elif time_start <= hr[i] - t0:
            first = i
        else:
            first = i + 1
        self.first = first

    @property
    def times(self):
        if self.first is None:
            return []
        return self.all_times[self.first:]

    def map(self, k):
        """Slices (3, ntm, npm1) of the k-th map from the first used one."""
        from . import h5io
        data, _ = h5io.read_map(os.path.join(self.root, self.files[self.first + k]))
        if data.ndim != 3 or data.shape[0] < 3:
            raise ValueError(f"assimilation map {self.files[self.first + k]}: shape {data.shape}")
        return data[:3]
