"""Minimal HDF5 writer/reader over the HDF5 C library (ctypes) for the CLI tests -- test
infrastructure only (h5py is not available).  Writes NLOS datasets in the field layout the
backprojector CLI reads and reads back its "voxelVolume" with its storage properties."""
from __future__ import annotations

import ctypes
import os

import numpy as np

H5_CANDIDATES = ["/opt/conda/lib/libhdf5.so", "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so",
                 "/usr/lib/x86_64-linux-gnu/libhdf5.so", "/usr/local/lib/libhdf5.so"]
hid_t = ctypes.c_int64
hsize_t = ctypes.c_uint64
_lib = None


def path():
    for p in H5_CANDIDATES:
        if os.path.exists(p):
            return p
    return None


def lib():
    global _lib
    if _lib is None:
        p = path()
        if p is None:
            raise RuntimeError("no HDF5 C library found")
        L = ctypes.CDLL(p)
        L.H5open()
        for name in ("H5Fcreate", "H5Fopen", "H5Screate_simple", "H5Screate", "H5Dcreate2", "H5Dopen2",
                     "H5Dget_space", "H5Dget_create_plist", "H5Tcopy", "H5Acreate2", "H5Pcreate"):
            getattr(L, name).restype = hid_t
        L.H5Fcreate.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]
        L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid_t]
        L.H5Screate_simple.argtypes = [ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]
        L.H5Screate.argtypes = [ctypes.c_int]
        L.H5Dcreate2.argtypes = [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]
        L.H5Dopen2.argtypes = [hid_t, ctypes.c_char_p, hid_t]
        L.H5Dwrite.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]
        L.H5Dread.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]
        L.H5Dget_space.argtypes = [hid_t]
        L.H5Dget_create_plist.argtypes = [hid_t]
        L.H5Sget_simple_extent_ndims.argtypes = [hid_t]
        L.H5Sget_simple_extent_dims.argtypes = [hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]
        L.H5Tcopy.argtypes = [hid_t]
        L.H5Tset_size.argtypes = [hid_t, ctypes.c_size_t]
        L.H5Acreate2.argtypes = [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t]
        L.H5Awrite.argtypes = [hid_t, hid_t, ctypes.c_void_p]
        L.H5Pget_chunk.argtypes = [hid_t, ctypes.c_int, ctypes.POINTER(hsize_t)]
        L.H5Pget_fill_value.argtypes = [hid_t, hid_t, ctypes.c_void_p]
        L.H5Pget_nfilters.argtypes = [hid_t]
        L.H5Pget_filter2.argtypes = [hid_t, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint),
                                     ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_uint),
                                     ctypes.c_size_t, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint)]
        for name in ("H5Fclose", "H5Sclose", "H5Dclose", "H5Aclose", "H5Tclose", "H5Pclose"):
            getattr(L, name).argtypes = [hid_t]
        _lib = L
    return _lib


def _type(name):
    return hid_t.in_dll(lib(), name).value


def _check(v, what):
    if v < 0:
        raise RuntimeError(f"HDF5 call failed: {what}")
    return v


def write(filename, fields, file_attrs=None, data_attrs=None):
    """fields: name -> array (float32 or int32); file_attrs: name -> str;
    data_attrs: name -> int (attributes of the "data" field)."""
    L = lib()
    f = _check(L.H5Fcreate(filename.encode(), 2, 0, 0), "H5Fcreate")  # H5F_ACC_TRUNC
    try:
        for name, arr in fields.items():
            a = np.ascontiguousarray(arr)
            t = _type("H5T_NATIVE_INT32_g") if a.dtype.kind in "iu" else _type("H5T_NATIVE_FLOAT_g")
            a = a.astype(np.int32 if a.dtype.kind in "iu" else np.float32)
            dims = (hsize_t * max(a.ndim, 1))(*(a.shape or (1,)))
            sp = _check(L.H5Screate_simple(max(a.ndim, 1), dims, None), "H5Screate_simple")
            ds = _check(L.H5Dcreate2(f, name.encode(), t, sp, 0, 0, 0), f"H5Dcreate2 {name}")
            _check(L.H5Dwrite(ds, t, 0, 0, 0, a.ctypes.data), f"H5Dwrite {name}")
            if name == "data":
                for an, av in (data_attrs or {}).items():
                    asp = L.H5Screate(0)  # scalar
                    at = L.H5Acreate2(ds, an.encode(), _type("H5T_NATIVE_INT_g"), asp, 0, 0)
                    v = ctypes.c_int(int(av))
                    _check(L.H5Awrite(at, _type("H5T_NATIVE_INT_g"), ctypes.byref(v)), "H5Awrite")
                    L.H5Aclose(at)
                    L.H5Sclose(asp)
            L.H5Dclose(ds)
            L.H5Sclose(sp)
        for an, av in (file_attrs or {}).items():
            b = av.encode()
            st = L.H5Tcopy(_type("H5T_C_S1_g"))
            L.H5Tset_size(st, len(b) + 1)
            asp = L.H5Screate(0)
            at = _check(L.H5Acreate2(f, an.encode(), st, asp, 0, 0), "H5Acreate2")
            buf = ctypes.create_string_buffer(b, len(b) + 1)
            _check(L.H5Awrite(at, st, buf), "H5Awrite")
            L.H5Aclose(at)
            L.H5Sclose(asp)
            L.H5Tclose(st)
    finally:
        L.H5Fclose(f)


def read_volume(filename, name="voxelVolume"):
    """(array float32, dict(chunk, deflate, fill))."""
    L = lib()
    f = _check(L.H5Fopen(filename.encode(), 0, 0), "H5Fopen")
    try:
        ds = _check(L.H5Dopen2(f, name.encode(), 0), f"H5Dopen2 {name}")
        sp = L.H5Dget_space(ds)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * nd)()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        out = np.zeros(tuple(int(d) for d in dims), np.float32)
        _check(L.H5Dread(ds, _type("H5T_NATIVE_FLOAT_g"), 0, 0, 0, out.ctypes.data), "H5Dread")
        pl = L.H5Dget_create_plist(ds)
        cdims = (hsize_t * nd)()
        L.H5Pget_chunk(pl, nd, cdims)
        fill = ctypes.c_float()
        L.H5Pget_fill_value(pl, _type("H5T_NATIVE_FLOAT_g"), ctypes.byref(fill))
        deflate = None
        for i in range(L.H5Pget_nfilters(pl)):
            flags, n = ctypes.c_uint(), ctypes.c_size_t(4)
            vals = (ctypes.c_uint * 4)()
            nm = ctypes.create_string_buffer(64)
            cfg = ctypes.c_uint()
            fid = L.H5Pget_filter2(pl, i, ctypes.byref(flags), ctypes.byref(n), vals, 64, nm, ctypes.byref(cfg))
            if fid == 1:  # H5Z_FILTER_DEFLATE
                deflate = int(vals[0])
        L.H5Pclose(pl)
        L.H5Sclose(sp)
        L.H5Dclose(ds)
        return out, dict(chunk=tuple(int(c) for c in cdims), deflate=deflate, fill=fill.value)
    finally:
        L.H5Fclose(f)


def nlos_dataset(filename, transient, camera_grid, laser_grid, camera_position, laser_position, t0,
                 deltaT, confocal, volume_position, volume_size, order="row-major", third_bounce=None):
    """Writes an NLOS dataset.  transient: row-major [lx][ly][(cx][cy])[bounce][T][channel];
    grids [X][Y][3].  order "column-major" stores every array with its axes reversed."""
    def o(a):
        a = np.asarray(a, np.float32)
        return a.T.copy() if order == "column-major" else a
    T = transient.shape[-2]
    fields = {
        "data": o(transient), "cameraGridPositions": o(camera_grid), "laserGridPositions": o(laser_grid),
        "cameraGridNormals": o(np.zeros_like(camera_grid)), "laserGridNormals": o(np.zeros_like(laser_grid)),
        "cameraPosition": np.asarray(camera_position, np.float32),
        "laserPosition": np.asarray(laser_position, np.float32),
        "cameraGridSize": np.array([2.0, 2.0], np.float32), "laserGridSize": np.array([2.0, 2.0], np.float32),
        "cameraGridPoints": np.array(camera_grid.shape[:2], np.float32),
        "laserGridPoints": np.array(laser_grid.shape[:2], np.float32),
        "hiddenVolumePosition": np.asarray(volume_position, np.float32),
        "hiddenVolumeRotation": np.zeros(3, np.float32),
        "hiddenVolumeSize": np.asarray(volume_size, np.float32),
        "t0": np.array([t0], np.float32), "deltaT": np.array([deltaT], np.float32),
        "t": np.array([T], np.int32), "isConfocal": np.array([1 if confocal else 0], np.int32),
    }
    attrs = {"data order": order} if order else {}
    write(filename, fields, attrs, {"third_bounce": third_bounce} if third_bounce is not None else None)
