"""A small HDF5 reader over the HDF5 C library (ctypes), for checking the
snapshot files the way the reference's reader does (python/runfiles.py:30-66:
datasets as arrays, file attributes as numbers).  The data goes back through
HDF5's own filter pipeline (inflate, unshuffle), so a file that reads back
right was encoded the way HDF5 expects."""
import ctypes
import os

import numpy as np

_CANDIDATES = ("libhdf5.so", "libhdf5_serial.so", "/opt/conda/lib/libhdf5.so",
               "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so", "/usr/local/lib/libhdf5.so")
_L = None

hid_t = ctypes.c_int64
hsize_t = ctypes.c_ulonglong


class _GInfo(ctypes.Structure):
    _fields_ = [("storage_type", ctypes.c_int), ("nlinks", hsize_t), ("max_corder", ctypes.c_int64),
                ("mounted", ctypes.c_bool)]


def lib():
    """The loaded HDF5 library, or None when it is absent."""
    global _L
    if _L is not None:
        return _L or None
    cands = ([os.environ["CB_HDF5_LIBRARY"]] if os.environ.get("CB_HDF5_LIBRARY") else []) + list(_CANDIDATES)
    for c in cands:
        try:
            L = ctypes.CDLL(c)
            break
        except OSError:
            continue
    else:
        _L = False
        return None
    L.H5open()
    L.H5Fopen.restype = hid_t
    L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid_t]
    L.H5Dopen2.restype = hid_t
    L.H5Dopen2.argtypes = [hid_t, ctypes.c_char_p, hid_t]
    L.H5Dget_space.restype = hid_t
    L.H5Dget_space.argtypes = [hid_t]
    L.H5Dget_type.restype = hid_t
    L.H5Dget_type.argtypes = [hid_t]
    L.H5Dget_create_plist.restype = hid_t
    L.H5Dget_create_plist.argtypes = [hid_t]
    L.H5Sget_simple_extent_ndims.argtypes = [hid_t]
    L.H5Sget_simple_extent_dims.argtypes = [hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]
    L.H5Dread.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]
    L.H5Aopen.restype = hid_t
    L.H5Aopen.argtypes = [hid_t, ctypes.c_char_p, hid_t]
    L.H5Aread.argtypes = [hid_t, hid_t, ctypes.c_void_p]
    L.H5Aget_num_attrs.argtypes = [hid_t]
    L.H5Aget_name_by_idx.restype = ctypes.c_ssize_t
    L.H5Aget_name_by_idx.argtypes = [hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, hsize_t,
                                     ctypes.c_char_p, ctypes.c_size_t, hid_t]
    L.H5Gget_info.argtypes = [hid_t, ctypes.POINTER(_GInfo)]
    L.H5Lget_name_by_idx.restype = ctypes.c_ssize_t
    L.H5Lget_name_by_idx.argtypes = [hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, hsize_t,
                                     ctypes.c_char_p, ctypes.c_size_t, hid_t]
    L.H5Pget_chunk.argtypes = [hid_t, ctypes.c_int, ctypes.POINTER(hsize_t)]
    L.H5Pget_nfilters.argtypes = [hid_t]
    L.H5Pget_filter2.argtypes = [hid_t, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint),
                                 ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_uint),
                                 ctypes.c_size_t, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint)]
    L.H5Tget_size.restype = ctypes.c_size_t
    L.H5Tget_size.argtypes = [hid_t]
    L.H5Tget_class.argtypes = [hid_t]
    L.H5Gopen2.restype = hid_t
    L.H5Gopen2.argtypes = [hid_t, ctypes.c_char_p, hid_t]
    for name in ("H5Fclose", "H5Dclose", "H5Sclose", "H5Aclose", "H5Pclose", "H5Tclose", "H5Gclose"):
        getattr(L, name).argtypes = [hid_t]
    _L = L
    return L


def _global(name):
    return hid_t.in_dll(lib(), name).value


class File:
    """One HDF5 file opened read-only."""

    def __init__(self, path):
        L = lib()
        self.f = L.H5Fopen(os.fsencode(path), 0, 0)       # H5F_ACC_RDONLY, H5P_DEFAULT
        if self.f < 0:
            raise OSError("cannot open " + path)

    def close(self):
        if self.f >= 0:
            lib().H5Fclose(self.f)
            self.f = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def names(self):
        """Link names of the root group, in name order."""
        L = lib()
        info = _GInfo()
        assert L.H5Gget_info(self.f, ctypes.byref(info)) >= 0
        out = []
        for i in range(info.nlinks):
            buf = ctypes.create_string_buffer(256)
            L.H5Lget_name_by_idx(self.f, b".", 0, 0, i, buf, 256, 0)
            out.append(buf.value.decode())
        return out

    def attrs(self):
        """Root-group attributes as {name: float}."""
        L = lib()
        out = {}
        root = L.H5Gopen2(self.f, b"/", 0)
        for i in range(L.H5Aget_num_attrs(root)):
            buf = ctypes.create_string_buffer(256)
            L.H5Aget_name_by_idx(root, b".", 0, 0, i, buf, 256, 0)
            a = L.H5Aopen(root, buf.value, 0)
            v = ctypes.c_double()
            assert L.H5Aread(a, _global("H5T_NATIVE_DOUBLE_g"), ctypes.byref(v)) >= 0
            L.H5Aclose(a)
            out[buf.value.decode()] = v.value
        L.H5Gclose(root)
        return out

    def dataset(self, name):
        """(values as float32 array, element bytes, chunk shape, filters) where
        filters is a list of (filter id, client values)."""
        L = lib()
        d = L.H5Dopen2(self.f, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        sp = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * nd)()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        L.H5Sclose(sp)
        ty = L.H5Dget_type(d)
        esize, tclass = L.H5Tget_size(ty), L.H5Tget_class(ty)
        L.H5Tclose(ty)
        out = np.empty(tuple(int(x) for x in dims), dtype=np.float32)
        assert L.H5Dread(d, _global("H5T_NATIVE_FLOAT_g"), 0, 0, 0, out.ctypes.data) >= 0
        pl = L.H5Dget_create_plist(d)
        ch = (hsize_t * nd)()
        L.H5Pget_chunk(pl, nd, ch)
        filters = []
        for i in range(L.H5Pget_nfilters(pl)):
            flags, nel, cfg = ctypes.c_uint(), ctypes.c_size_t(8), ctypes.c_uint()
            cd = (ctypes.c_uint * 8)()
            nm = ctypes.create_string_buffer(64)
            fid = L.H5Pget_filter2(pl, i, ctypes.byref(flags), ctypes.byref(nel), cd, 64, nm, ctypes.byref(cfg))
            filters.append((fid, tuple(cd[:nel.value])))
        L.H5Pclose(pl)
        L.H5Dclose(d)
        return out, (esize, tclass), tuple(int(x) for x in ch), filters
