"""HDF5 files through the HDF5 C library (ctypes).

goSPL writes its outputs with h5py (gospl/tools/outmesh.py:256-507), which
wraps the HDF5 C library; the image's Python has no h5py, so this module
binds the library itself, with only what goSPL's files need: 2-D datasets of
float32 / int32 / float64 at the file root, optionally gzip-compressed
(h5py's ``compression="gzip"``), written and read whole.

The library is looked for in ``GOSPL_HDF5_LIB``, then as ``libhdf5.so`` on
the loader path, then under the Conda prefix the image ships
(``/opt/conda/lib``).  A missing library is an error at first use.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_hid = ctypes.c_int64        # hid_t (HDF5 >= 1.10)
_hsize = ctypes.c_uint64     # hsize_t
H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0

_LIB = None
_TYPES = {}


def _candidates():
    env = os.environ.get("GOSPL_HDF5_LIB")
    if env:
        yield env
    yield "libhdf5.so"
    for pre in (os.environ.get("CONDA_PREFIX"), "/opt/conda"):
        if pre:
            yield os.path.join(pre, "lib", "libhdf5.so")


def lib():
    """The loaded HDF5 library (opened once)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    err = []
    for cand in _candidates():
        try:
            L = ctypes.CDLL(cand)
            break
        except OSError as e:
            err.append("%s: %s" % (cand, e))
    else:
        raise RuntimeError("HDF5 output needs the HDF5 C library (set GOSPL_HDF5_LIB); tried " + "; ".join(err))
    for name, res, args in (
            ("H5open", ctypes.c_int, []),
            ("H5get_libversion", ctypes.c_int, [ctypes.POINTER(ctypes.c_uint)] * 3),
            ("H5Fcreate", _hid, [ctypes.c_char_p, ctypes.c_uint, _hid, _hid]),
            ("H5Fopen", _hid, [ctypes.c_char_p, ctypes.c_uint, _hid]),
            ("H5Fclose", ctypes.c_int, [_hid]),
            ("H5Screate_simple", _hid, [ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
            ("H5Sclose", ctypes.c_int, [_hid]),
            ("H5Sget_simple_extent_dims", ctypes.c_int, [_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
            ("H5Pcreate", _hid, [_hid]),
            ("H5Pset_chunk", ctypes.c_int, [_hid, ctypes.c_int, ctypes.POINTER(_hsize)]),
            ("H5Pset_deflate", ctypes.c_int, [_hid, ctypes.c_uint]),
            ("H5Pclose", ctypes.c_int, [_hid]),
            ("H5Dcreate2", _hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid]),
            ("H5Dopen2", _hid, [_hid, ctypes.c_char_p, _hid]),
            ("H5Dwrite", ctypes.c_int, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
            ("H5Dread", ctypes.c_int, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
            ("H5Dget_space", _hid, [_hid]),
            ("H5Dclose", ctypes.c_int, [_hid]),
            ("H5Lexists", ctypes.c_int, [_hid, ctypes.c_char_p, _hid])):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.H5open() < 0:
        raise RuntimeError("H5open failed")
    for key, sym in (("f4", "H5T_NATIVE_FLOAT_g"), ("f8", "H5T_NATIVE_DOUBLE_g"), ("i4", "H5T_NATIVE_INT_g"),
                     ("dcpl", "H5P_CLS_DATASET_CREATE_ID_g")):
        _TYPES[key] = _hid.in_dll(L, sym).value
    _LIB = L
    return L


def version():
    L = lib()
    v = [ctypes.c_uint() for _ in range(3)]
    L.H5get_libversion(*[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def _ok(rc, what):
    if rc < 0:
        raise RuntimeError("HDF5: %s failed" % what)
    return rc


def _dtype_key(a):
    k = {np.dtype(np.float32): "f4", np.dtype(np.float64): "f8", np.dtype(np.int32): "i4"}.get(a.dtype)
    if k is None:
        raise TypeError("HDF5 datasets here are float32, float64 or int32 (got %s)" % a.dtype)
    return k


class File:
    """An HDF5 file opened for writing (``mode="w"``, truncating, as h5py's
    ``File(name, "w")``) or reading (``"r"``)."""

    def __init__(self, path, mode="r"):
        L = lib()
        self.L = L
        self.path = str(path)
        if mode == "w":
            self.fid = _ok(L.H5Fcreate(self.path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
                           "create " + self.path)
        elif mode == "r":
            if not os.path.exists(self.path):
                raise FileNotFoundError(self.path)
            self.fid = _ok(L.H5Fopen(self.path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "open " + self.path)
        else:
            raise ValueError("mode must be 'w' or 'r'")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if getattr(self, "fid", None) is not None:
            self.L.H5Fclose(self.fid)
            self.fid = None

    def write(self, name, data, compression="gzip", level=4):
        """Create dataset ``name`` holding ``data`` (2-D; C order).
        compression None writes it contiguous; "gzip" (or an int level) as
        chunked deflate, h5py's default level being 4."""
        L = self.L
        a = np.ascontiguousarray(data)
        if a.ndim != 2:
            raise ValueError("datasets here are 2-D (got shape %s)" % (a.shape,))
        k = _dtype_key(a)
        dims = (_hsize * 2)(*a.shape)
        sid = _ok(L.H5Screate_simple(2, dims, None), "dataspace")
        pid = H5P_DEFAULT
        try:
            if isinstance(compression, bool):
                compression = "gzip" if compression else None
            if isinstance(compression, int):
                level, compression = compression, "gzip"
            if compression is not None and a.size > 0:
                if compression != "gzip":
                    raise ValueError("only gzip compression is supported (got %r)" % compression)
                pid = _ok(L.H5Pcreate(_TYPES["dcpl"]), "property list")
                rows = max(1, min(a.shape[0], (1 << 20) // max(1, a.shape[1] * a.itemsize)))
                chunk = (_hsize * 2)(rows, a.shape[1])
                _ok(L.H5Pset_chunk(pid, 2, chunk), "chunk")
                _ok(L.H5Pset_deflate(pid, int(level)), "deflate")
            did = _ok(L.H5Dcreate2(self.fid, name.encode(), _TYPES[k], sid, H5P_DEFAULT, pid, H5P_DEFAULT),
                      "create dataset " + name)
            try:
                if a.size:
                    _ok(L.H5Dwrite(did, _TYPES[k], H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data),
                        "write " + name)
            finally:
                L.H5Dclose(did)
        finally:
            if pid != H5P_DEFAULT:
                L.H5Pclose(pid)
            L.H5Sclose(sid)

    def __contains__(self, name):
        n = name.lstrip("/").encode()
        return self.L.H5Lexists(self.fid, n, H5P_DEFAULT) > 0

    def read(self, name, dtype=np.float64):
        """Dataset ``name`` as a 2-D array of ``dtype`` (HDF5 converts)."""
        L = self.L
        did = _ok(L.H5Dopen2(self.fid, name.lstrip("/").encode(), H5P_DEFAULT), "open dataset " + name)
        try:
            sid = _ok(L.H5Dget_space(did), "dataspace")
            dims = (_hsize * 2)()
            nd = _ok(L.H5Sget_simple_extent_dims(sid, dims, None), "dims")
            L.H5Sclose(sid)
            shape = tuple(int(dims[i]) for i in range(nd))
            out = np.empty(shape, dtype=dtype)
            k = _dtype_key(out)
            if out.size:
                _ok(L.H5Dread(did, _TYPES[k], H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data), "read " + name)
            return out
        finally:
            L.H5Dclose(did)
