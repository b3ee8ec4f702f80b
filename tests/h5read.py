"""Read whole datasets of an HDF5 file into numpy (test helper; h5py is not
installed).  Uses the image's HDF5 runtime through ctypes: open the file
read-only, read a dataset's extent, read it as native doubles."""
import ctypes
import os

import numpy as np

HDF5_LIB = os.path.join(os.environ.get("HDF5_PREFIX", "/opt/conda"), "lib", "libhdf5.so")

_lib = None


def _h5():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(HDF5_LIB)
        lib.H5open.restype = ctypes.c_int
        if lib.H5open() < 0:
            raise OSError("H5open failed")
        hid = ctypes.c_int64
        lib.H5Fopen.restype = hid
        lib.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid]
        lib.H5Fclose.argtypes = [hid]
        lib.H5Dopen2.restype = hid
        lib.H5Dopen2.argtypes = [hid, ctypes.c_char_p, hid]
        lib.H5Dclose.argtypes = [hid]
        lib.H5Dget_space.restype = hid
        lib.H5Dget_space.argtypes = [hid]
        lib.H5Sclose.argtypes = [hid]
        lib.H5Sget_simple_extent_ndims.argtypes = [hid]
        lib.H5Sget_simple_extent_dims.argtypes = [hid, ctypes.POINTER(ctypes.c_uint64),
                                                  ctypes.c_void_p]
        lib.H5Dread.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
        lib.H5Lexists.restype = ctypes.c_int
        lib.H5Lexists.argtypes = [hid, ctypes.c_char_p, hid]
        _lib = lib
    return _lib


def available():
    try:
        _h5()
        return True
    except OSError:
        return False


def read(path, name):
    """Dataset `name` of file `path` as a float64 array of its own shape
    (None if the file has no such dataset)."""
    lib = _h5()
    f = lib.H5Fopen(path.encode(), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
    if f < 0:
        raise OSError(f"cannot open {path}")
    try:
        if lib.H5Lexists(f, name.encode(), 0) <= 0:
            return None
        d = lib.H5Dopen2(f, name.encode(), 0)
        if d < 0:
            raise OSError(f"cannot open {name} in {path}")
        try:
            sp = lib.H5Dget_space(d)
            nd = lib.H5Sget_simple_extent_ndims(sp)
            dims = (ctypes.c_uint64 * max(nd, 1))()
            lib.H5Sget_simple_extent_dims(sp, dims, None)
            lib.H5Sclose(sp)
            shape = tuple(int(dims[q]) for q in range(nd))
            out = np.empty(shape, dtype=np.float64)
            native_double = ctypes.c_int64.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
            rc = lib.H5Dread(d, native_double, 0, 0, 0, out.ctypes.data_as(ctypes.c_void_p))
            if rc < 0:
                raise OSError(f"cannot read {name} in {path}")
            return out
        finally:
            lib.H5Dclose(d)
    finally:
        lib.H5Fclose(f)
