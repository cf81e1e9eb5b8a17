"""HDF5 map files in HipFT's format, through the system HDF5 C library.

A map file holds the dataset ``Data`` and its dimension scales ``dim1``,
``dim2`` (and ``dim3``), written by the reference's wrh5
(src/hipft.f90:6155-6351) from Fortran arrays f(n1,n2[,n3]): phi (n1 =
npm-1 points, one-point overlap) fastest, then theta (n2 = ntm), then the
realization index (n3, scale 1..nr).  Seen from C the dataspace is
(n3,) n2, n1 and dimension scale ``dimI`` is attached to C dimension
rank - I.  ``Data`` and the scales are float32 unless double precision is
asked for (``output_single_precision``, 548).  Reading accepts either
precision (rdh5, 5678-6019).

read_map(path)  -> (data, scales): data float64 with shape (n2, n1) or
                   (n3, n2, n1); scales [s1 (phi), s2 (theta), (s3)].
write_map(path, data, scales, single=True)
"""
import ctypes
import os

import numpy as np

_LIB = None
hid_t = ctypes.c_int64
hsize_t = ctypes.c_uint64
herr_t = ctypes.c_int

H5F_ACC_RDONLY = 0
H5F_ACC_TRUNC = 2
H5P_DEFAULT = 0
H5S_ALL = 0


class H5Error(RuntimeError):
    pass


def _lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    paths = [os.environ.get("HFT_HDF5_LIB"), "/opt/conda/lib/libhdf5.so", "libhdf5.so"]
    hl_paths = [os.environ.get("HFT_HDF5_HL_LIB"), "/opt/conda/lib/libhdf5_hl.so",
                "libhdf5_hl.so"]
    h5 = hl = None
    for p in paths:
        if p:
            try:
                h5 = ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
                break
            except OSError:
                continue
    for p in hl_paths:
        if p:
            try:
                hl = ctypes.CDLL(p)
                break
            except OSError:
                continue
    if h5 is None or hl is None:
        raise H5Error("the HDF5 C library (libhdf5, libhdf5_hl) is not available")
    sig = {
        "H5open": (herr_t, []),
        "H5Fopen": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]),
        "H5Fcreate": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]),
        "H5Fclose": (herr_t, [hid_t]),
        "H5Dopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Dcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dget_space": (hid_t, [hid_t]),
        "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dclose": (herr_t, [hid_t]),
        "H5Screate_simple": (hid_t, [ctypes.c_int, ctypes.POINTER(hsize_t),
                                     ctypes.POINTER(hsize_t)]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid_t]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [hid_t, ctypes.POINTER(hsize_t),
                                                     ctypes.POINTER(hsize_t)]),
        "H5Sclose": (herr_t, [hid_t]),
        "H5Lexists": (ctypes.c_int, [hid_t, ctypes.c_char_p, hid_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(h5, name)
        fn.restype, fn.argtypes = res, args
    for name, args in (("H5DSset_scale", [hid_t, ctypes.c_char_p]),
                       ("H5DSattach_scale", [hid_t, hid_t, ctypes.c_uint]),
                       ("H5DSset_label", [hid_t, ctypes.c_uint, ctypes.c_char_p])):
        fn = getattr(hl, name)
        fn.restype, fn.argtypes = herr_t, args
    h5.H5open()
    _LIB = (h5, hl)
    return _LIB


def _native(h5, name):
    return hid_t.in_dll(h5, name).value


def _check(v, what):
    if v < 0:
        raise H5Error(f"HDF5 call {what} failed ({v})")
    return v


def _read_dataset(h5, fid, name):
    did = _check(h5.H5Dopen2(fid, name.encode(), H5P_DEFAULT), f"H5Dopen2({name})")
    try:
        sid = _check(h5.H5Dget_space(did), "H5Dget_space")
        nd = h5.H5Sget_simple_extent_ndims(sid)
        dims = (hsize_t * max(nd, 1))()
        h5.H5Sget_simple_extent_dims(sid, dims, None)
        h5.H5Sclose(sid)
        shape = tuple(int(dims[i]) for i in range(nd))
        out = np.empty(shape, dtype=np.float64)
        _check(h5.H5Dread(did, _native(h5, "H5T_NATIVE_DOUBLE_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT,
                          out.ctypes.data), f"H5Dread({name})")
        return out
    finally:
        h5.H5Dclose(did)


def read_map(path):
    """Read a HipFT map file: (data float64 (n2,n1) or (n3,n2,n1), [s1, s2, (s3)])."""
    h5, _ = _lib()
    if not os.path.exists(path):
        raise H5Error(f"map file not found: {path}")
    fid = _check(h5.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, H5P_DEFAULT), f"H5Fopen({path})")
    try:
        data = _read_dataset(h5, fid, "Data")
        scales = []
        for i in range(1, data.ndim + 1):
            name = f"dim{i}"
            if h5.H5Lexists(fid, name.encode(), H5P_DEFAULT) > 0:
                scales.append(_read_dataset(h5, fid, name))
            else:
                scales.append(np.arange(1, data.shape[data.ndim - i] + 1, dtype=np.float64))
        return data, scales
    finally:
        h5.H5Fclose(fid)


def write_map(path, data, scales, single=True):
    """Write a HipFT map file (``Data`` + dimension scales ``dim1..``)."""
    h5, hl = _lib()
    data = np.ascontiguousarray(data, dtype=np.float64)
    rank = data.ndim
    if rank not in (2, 3) or len(scales) != rank:
        raise H5Error("write_map: data must be 2-D or 3-D with one scale per dimension")
    for i, sc in enumerate(scales):
        if len(sc) != data.shape[rank - 1 - i]:
            raise H5Error(f"write_map: scale dim{i + 1} has {len(sc)} values, "
                          f"the data {data.shape[rank - 1 - i]}")
    ftype = _native(h5, "H5T_NATIVE_FLOAT_g" if single else "H5T_NATIVE_DOUBLE_g")
    mtype = _native(h5, "H5T_NATIVE_DOUBLE_g")
    fid = _check(h5.H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
                 f"H5Fcreate({path})")
    try:
        dims = (hsize_t * rank)(*data.shape)
        sid = _check(h5.H5Screate_simple(rank, dims, None), "H5Screate_simple")
        did = _check(h5.H5Dcreate2(fid, b"Data", ftype, sid, H5P_DEFAULT, H5P_DEFAULT,
                                   H5P_DEFAULT), "H5Dcreate2(Data)")
        try:
            _check(h5.H5Dwrite(did, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data.ctypes.data),
                   "H5Dwrite(Data)")
            for i, sc in enumerate(scales):
                name = f"dim{i + 1}".encode()
                v = np.ascontiguousarray(sc, dtype=np.float64)
                d1 = (hsize_t * 1)(len(v))
                ssid = _check(h5.H5Screate_simple(1, d1, None), "H5Screate_simple")
                sdid = _check(h5.H5Dcreate2(fid, name, ftype, ssid, H5P_DEFAULT, H5P_DEFAULT,
                                            H5P_DEFAULT), "H5Dcreate2(scale)")
                _check(h5.H5Dwrite(sdid, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.ctypes.data),
                       "H5Dwrite(scale)")
                cdim = rank - 1 - i
                _check(hl.H5DSset_scale(sdid, name), "H5DSset_scale")
                _check(hl.H5DSattach_scale(did, sdid, cdim), "H5DSattach_scale")
                _check(hl.H5DSset_label(did, cdim, name), "H5DSset_label")
                h5.H5Dclose(sdid)
                h5.H5Sclose(ssid)
        finally:
            h5.H5Dclose(did)
            h5.H5Sclose(sid)
    finally:
        h5.H5Fclose(fid)


def map_from_state(model, pt_state, t, p, single=True, path=None):
    """PT map(s) of a FluxTransport state in the reference's output layout
    (write_map, src/hipft.f90:2801-2917): 2-D for one realization, else 3-D
    with realization scale 1..nr.  Returns (data, scales); writes when path."""
    nr = pt_state.shape[0]
    if nr == 1:
        data, scales = pt_state[0], [p, t]
    else:
        data, scales = pt_state, [p, t, np.arange(1, nr + 1, dtype=np.float64)]
    if path:
        write_map(path, data, scales, single=single)
    return data, scales
