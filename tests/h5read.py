"""A minimal HDF5 reader for the tests (ctypes over the HDF5 C library), independent of the
product's writer: it lists the root group and reads scalar string datasets and double
datasets, so the tests can check .pyfrs files against the layout the reference's reader
expects (src/solver.cpp:4670-4837)."""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_lib = None
hid_t = C.c_int64
hsize_t = C.c_ulonglong


def available():
    try:
        lib()
        return True
    except OSError:
        return False


def lib():
    global _lib
    if _lib is None:
        names = [os.environ.get("ZEFR_HDF5_LIBRARY"), "libhdf5.so", "/opt/conda/lib/libhdf5.so"]
        err = None
        for n in names:
            if not n:
                continue
            try:
                L = C.CDLL(n)
                break
            except OSError as e:
                err = e
        else:
            raise OSError(f"no HDF5 library: {err}")
        L.H5open()
        L.H5Eset_auto2(hid_t(0), None, None)
        L.H5Fopen.restype = hid_t
        L.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid_t]
        L.H5Fclose.argtypes = [hid_t]
        L.H5Dopen2.restype = hid_t
        L.H5Dopen2.argtypes = [hid_t, C.c_char_p, hid_t]
        L.H5Dclose.argtypes = [hid_t]
        L.H5Dget_space.restype = hid_t
        L.H5Dget_space.argtypes = [hid_t]
        L.H5Dget_type.restype = hid_t
        L.H5Dget_type.argtypes = [hid_t]
        L.H5Sget_simple_extent_ndims.argtypes = [hid_t]
        L.H5Sget_simple_extent_dims.argtypes = [hid_t, C.POINTER(hsize_t), C.c_void_p]
        L.H5Sclose.argtypes = [hid_t]
        L.H5Tcopy.restype = hid_t
        L.H5Tcopy.argtypes = [hid_t]
        L.H5Tset_size.argtypes = [hid_t, C.c_size_t]
        L.H5Tis_variable_str.argtypes = [hid_t]
        L.H5Tget_class.argtypes = [hid_t]
        L.H5Tclose.argtypes = [hid_t]
        L.H5Dread.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]
        L.H5free_memory.argtypes = [C.c_void_p]
        L.H5Gget_info.argtypes = [hid_t, C.c_void_p]
        L.H5Lget_name_by_idx.restype = C.c_ssize_t
        L.H5Lget_name_by_idx.argtypes = [hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p, C.c_size_t,
                                         hid_t]
        _lib = L
    return _lib


def _gid(name):
    return hid_t.in_dll(lib(), name).value


class File:
    def __init__(self, path):
        self.h = lib().H5Fopen(path.encode(), 0, 0)
        if self.h < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        if self.h >= 0:
            lib().H5Fclose(self.h)
            self.h = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def names(self):
        """Names of the links in the root group, in name order."""
        # H5G_info_t: storage_type (int enum), nlinks (hsize_t), max_corder (int64), mounted (hbool_t)
        class GInfo(C.Structure):
            _fields_ = [("storage_type", C.c_int), ("nlinks", hsize_t), ("max_corder", C.c_int64),
                        ("mounted", C.c_bool)]
        info = GInfo()
        assert lib().H5Gget_info(self.h, C.byref(info)) >= 0
        out = []
        for i in range(info.nlinks):
            buf = C.create_string_buffer(256)
            # H5_INDEX_NAME = 0, H5_ITER_INC = 0
            n = lib().H5Lget_name_by_idx(self.h, b".", 0, 0, i, buf, 256, 0)
            assert n >= 0
            out.append(buf.value.decode())
        return out

    def _open(self, name):
        d = lib().H5Dopen2(self.h, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        return d

    def string(self, name):
        L = lib()
        d = self._open(name)
        ft = L.H5Dget_type(d)
        assert L.H5Tget_class(ft) == 3, "not a string dataset"  # H5T_STRING
        variable = L.H5Tis_variable_str(ft) > 0
        mt = L.H5Tcopy(_gid("H5T_C_S1_g"))
        L.H5Tset_size(mt, C.c_size_t(-1).value)
        p = C.c_char_p()
        assert L.H5Dread(d, mt, 0, 0, 0, C.byref(p)) >= 0
        s = p.value.decode()
        L.H5free_memory(p)
        L.H5Tclose(mt)
        L.H5Tclose(ft)
        L.H5Dclose(d)
        return s, variable

    def doubles(self, name):
        L = lib()
        d = self._open(name)
        sp = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * nd)()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        L.H5Sclose(sp)
        a = np.empty(tuple(int(x) for x in dims), np.float64)
        assert L.H5Dread(d, _gid("H5T_NATIVE_DOUBLE_g"), 0, 0, 0, a.ctypes.data) >= 0
        L.H5Dclose(d)
        return a
