#!/usr/bin/env python3
"""Write the reference-testsuite golden fixtures used by the parity tests.

This script reads DATA files that the reference's own regression suite holds
(``testsuite/<case>/reference/hipft_history_{sol,num}_r*.out``, the text
outputs ``hipft_realization_parameters.out`` and ``hipft_timing.out``, and the
HDF5 input maps ``testsuite/<case>/input/*.h5``) and stores them as compact
numpy archives under ``tests/golden/``.  Nothing from the reference is imported or
executed: the history files are parsed as text and the maps are read through
the system HDF5 C library (``libhdf5.so``) with ``ctypes``.

Run once in a container where the reference tree is mounted:

    python tests/golden/make_fixtures.py /root/reference
    python tests/golden/make_fixtures.py /root/reference --run-parameters
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = [
    "advect_gaussians_phi",
    "advect_gaussians_theta",
    "advect_gaussians_phi_theta",
    "diffuse_soccer",
    "diffuse_advect_soccer",
    "diffuse_dipole",
    "diffuse_advect_atten_map_1cr",
]


def read_history(path):
    """Parse a hipft_history_*.out file into a float64 array (rows x cols)."""
    rows = []
    with open(path) as fh:
        header = fh.readline().split()
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            rows.append([float(x.replace("D", "E")) for x in parts])
    return header, np.array(rows, dtype=np.float64)


def _hdf5():
    for cand in ("/opt/conda/lib/libhdf5.so", "libhdf5.so"):
        try:
            return ctypes.CDLL(cand)
        except OSError:
            continue
    raise RuntimeError("libhdf5.so not found")


def read_h5_float32(path, name):
    """Read one float32 dataset of an HDF5 file with the HDF5 C API."""
    h5 = _hdf5()
    h5.H5open()
    hid_t = ctypes.c_int64
    h5.H5Fopen.restype = hid_t
    h5.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid_t]
    h5.H5Dopen2.restype = hid_t
    h5.H5Dopen2.argtypes = [hid_t, ctypes.c_char_p, hid_t]
    h5.H5Dget_space.restype = hid_t
    h5.H5Dget_space.argtypes = [hid_t]
    h5.H5Sget_simple_extent_ndims.argtypes = [hid_t]
    h5.H5Sget_simple_extent_dims.argtypes = [hid_t, ctypes.POINTER(ctypes.c_uint64),
                                             ctypes.POINTER(ctypes.c_uint64)]
    h5.H5Dread.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]
    h5.H5Dclose.argtypes = [hid_t]
    h5.H5Sclose.argtypes = [hid_t]
    h5.H5Fclose.argtypes = [hid_t]
    native_float = hid_t.in_dll(h5, "H5T_NATIVE_FLOAT_g").value
    fid = h5.H5Fopen(path.encode(), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
    if fid < 0:
        raise RuntimeError(f"cannot open {path}")
    did = h5.H5Dopen2(fid, name.encode(), 0)
    sid = h5.H5Dget_space(did)
    nd = h5.H5Sget_simple_extent_ndims(sid)
    dims = (ctypes.c_uint64 * nd)()
    h5.H5Sget_simple_extent_dims(sid, dims, None)
    shape = tuple(int(d) for d in dims)
    out = np.empty(shape, dtype=np.float32)
    rc = h5.H5Dread(did, native_float, 0, 0, 0, out.ctypes.data)  # H5S_ALL, H5P_DEFAULT
    h5.H5Sclose(sid)
    h5.H5Dclose(did)
    h5.H5Fclose(fid)
    if rc < 0:
        raise RuntimeError(f"H5Dread failed for {path}:{name}")
    return out


def main(ref_root):
    ts = os.path.join(ref_root, "testsuite")
    for case in CASES:
        refdir = os.path.join(ts, case, "reference")
        sol, num = [], []
        r = 1
        while os.path.exists(os.path.join(refdir, f"hipft_history_sol_r{r:06d}.out")):
            _, s = read_history(os.path.join(refdir, f"hipft_history_sol_r{r:06d}.out"))
            _, n = read_history(os.path.join(refdir, f"hipft_history_num_r{r:06d}.out"))
            sol.append(s)
            num.append(n)
            r += 1
        arrays = {"sol": np.stack(sol), "num": np.stack(num)}
        indir = os.path.join(ts, case, "input")
        h5s = [f for f in sorted(os.listdir(indir)) if f.endswith(".h5")]
        if h5s:
            path = os.path.join(indir, h5s[0])
            # HDF5 stores C order (nt, np); the reference reads it as Fortran (np, nt) = PT.
            arrays["map_pt"] = read_h5_float32(path, "Data")
            arrays["scale_p"] = read_h5_float32(path, "dim1")
            arrays["scale_t"] = read_h5_float32(path, "dim2")
        with open(os.path.join(indir, "hipft.in")) as fh:
            arrays["hipft_in"] = np.array(fh.read())
        for name, key in (("hipft_realization_parameters.out", "realization_parameters"),
                          ("hipft_timing.out", "timing")):
            path = os.path.join(refdir, name)
            if os.path.exists(path):
                with open(path) as fh:
                    arrays[key] = np.array(fh.read())
        out = os.path.join(HERE, f"{case}.npz")
        np.savez_compressed(out, **arrays)
        print(case, {k: v.shape for k, v in arrays.items()}, os.path.getsize(out))


def run_parameters(ref_root, cases=("diffuse_soccer", "diffuse_advect_atten_map_1cr")):
    """hipft_run_parameters_used.out of two cases (the namelist values a run
    used, as the reference wrote them) into run_parameters_used.npz."""
    arrays = {}
    for case in cases:
        path = os.path.join(ref_root, "testsuite", case, "reference",
                            "hipft_run_parameters_used.out")
        with open(path) as fh:
            arrays[case] = np.array(fh.read())
    out = os.path.join(HERE, "run_parameters_used.npz")
    np.savez_compressed(out, **arrays)
    print("run_parameters_used", list(arrays), os.path.getsize(out))


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    if len(sys.argv) > 2 and sys.argv[2] == "--run-parameters":
        run_parameters(root)
    else:
        main(root)
