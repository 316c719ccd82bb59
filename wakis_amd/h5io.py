"""HDF5 files in the reference's layouts: ``Ez.h5`` and ``solver_state.h5``.

The reference writes and reads them with h5py (routines.py:376-445,
solverFIT3D.py:886-993, wakeSolver.py:1313-1360). This module gives the same
subset of h5py's interface that those call sites use:

    f = File(name, "w");  f["x"] = array;  f.create_dataset("E", data=array)
    f = File(name, "r");  f.keys();  "t" in f;  np.array(f["t"]);  f["E"][:]
    f["#00012"].shape;  f["#00012"][1, 1, 3:-3];  f.close()  (or ``with``)

Every dataset is a plain numeric array in the root group (scalars are
zero-dimensional datasets), as h5py stores a numpy array or scalar.

When h5py imports, ``File`` is ``h5py.File`` itself. Otherwise the HDF5 C
library is called directly through ctypes (HDF5 >= 1.10, where ``hid_t`` is 64
bits): the library named by ``WAKIS_HDF5_LIB``, else the one the dynamic
loader finds, else a libhdf5 of a conda or system install. Files written
either way are ordinary HDF5 files: h5py and h5dump read them, and this module
reads theirs.
"""

import ctypes
import ctypes.util
import glob
import os

import numpy as np

try:  # the reference's own dependency, when present
    import h5py as _h5py
except ImportError:  # pragma: no cover - depends on the install
    _h5py = None

__all__ = ["File", "backend", "is_hdf5"]

_SIGNATURE = b"\x89HDF\r\n\x1a\n"

_hid = ctypes.c_int64
_herr = ctypes.c_int
_hsize = ctypes.c_uint64

# H5Fpublic.h / H5Ppublic.h / H5Spublic.h / H5Tpublic.h / H5public.h constants
_ACC_RDONLY, _ACC_RDWR, _ACC_TRUNC = 0x0000, 0x0001, 0x0002
_P_DEFAULT = 0
_S_ALL = 0
_S_SCALAR = 0
_T_INTEGER, _T_FLOAT = 0, 1
_T_SGN_NONE = 0
_INDEX_NAME, _ITER_INC = 0, 0
_E_DEFAULT = 0

_LIB = None


def is_hdf5(path):
    """True if the file at ``path`` starts with the HDF5 signature."""
    with open(path, "rb") as f:
        return f.read(8) == _SIGNATURE


def _candidates():
    env = os.environ.get("WAKIS_HDF5_LIB")
    if env:
        yield env
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found
    for pat in ("/opt/conda/lib/libhdf5.so*", "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so*",
                "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so*", "/usr/lib/x86_64-linux-gnu/libhdf5.so*",
                "/usr/local/lib/libhdf5.so*", "/usr/lib64/libhdf5.so*"):
        yield from sorted(glob.glob(pat))


def _bind(lib):
    def f(name, res, *args):
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, list(args)
        return fn

    P = ctypes.POINTER
    f("H5open", _herr)
    f("H5get_libversion", _herr, P(ctypes.c_uint), P(ctypes.c_uint), P(ctypes.c_uint))
    f("H5Eset_auto2", _herr, _hid, ctypes.c_void_p, ctypes.c_void_p)
    f("H5Fcreate", _hid, ctypes.c_char_p, ctypes.c_uint, _hid, _hid)
    f("H5Fopen", _hid, ctypes.c_char_p, ctypes.c_uint, _hid)
    f("H5Fclose", _herr, _hid)
    f("H5Screate", _hid, ctypes.c_int)
    f("H5Screate_simple", _hid, ctypes.c_int, P(_hsize), P(_hsize))
    f("H5Sget_simple_extent_ndims", ctypes.c_int, _hid)
    f("H5Sget_simple_extent_dims", ctypes.c_int, _hid, P(_hsize), P(_hsize))
    f("H5Sclose", _herr, _hid)
    f("H5Dcreate2", _hid, _hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid)
    f("H5Dopen2", _hid, _hid, ctypes.c_char_p, _hid)
    f("H5Dwrite", _herr, _hid, _hid, _hid, _hid, _hid, ctypes.c_void_p)
    f("H5Dread", _herr, _hid, _hid, _hid, _hid, _hid, ctypes.c_void_p)
    f("H5Dget_space", _hid, _hid)
    f("H5Dget_type", _hid, _hid)
    f("H5Dclose", _herr, _hid)
    f("H5Tget_class", ctypes.c_int, _hid)
    f("H5Tget_size", ctypes.c_size_t, _hid)
    f("H5Tget_sign", ctypes.c_int, _hid)
    f("H5Tclose", _herr, _hid)
    f("H5Lexists", ctypes.c_int, _hid, ctypes.c_char_p, _hid)
    f("H5Ldelete", _herr, _hid, ctypes.c_char_p, _hid)
    f("H5Lget_name_by_idx", ctypes.c_ssize_t, _hid, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
      _hsize, ctypes.c_char_p, ctypes.c_size_t, _hid)
    f("H5Gget_info", _herr, _hid, ctypes.c_void_p)


def _lib():
    """The HDF5 C library, loaded and initialised once."""
    global _LIB
    if _LIB is not None:
        return _LIB
    errors = []
    for cand in _candidates():
        try:
            lib = ctypes.CDLL(cand)
            _bind(lib)
        except (OSError, AttributeError) as e:
            errors.append(f"{cand}: {e}")
            continue
        if lib.H5open() < 0:
            errors.append(f"{cand}: H5open failed")
            continue
        v = [ctypes.c_uint() for _ in range(3)]
        lib.H5get_libversion(*(ctypes.byref(x) for x in v))
        ver = tuple(x.value for x in v)
        if ver < (1, 10, 0):  # 32-bit hid_t before 1.10
            errors.append(f"{cand}: HDF5 {ver} (1.10 or later is needed)")
            continue
        lib.H5Eset_auto2(_E_DEFAULT, None, None)  # errors come back as status codes
        lib.version = ver
        lib.path = cand
        types = {}
        for name, dt in (("DOUBLE", "f8"), ("FLOAT", "f4"), ("INT8", "i1"), ("UINT8", "u1"),
                         ("INT16", "i2"), ("UINT16", "u2"), ("INT32", "i4"), ("UINT32", "u4"),
                         ("INT64", "i8"), ("UINT64", "u8")):
            types[np.dtype(dt)] = _hid.in_dll(lib, f"H5T_NATIVE_{name}_g").value
        lib.types = types
        _LIB = lib
        return lib
    raise ImportError("HDF5 files need h5py or the HDF5 C library (libhdf5 1.10+); none was "
                      "found" + (": " + "; ".join(errors) if errors else "") +
                      ". Set WAKIS_HDF5_LIB to the library's path.")


def backend():
    """'h5py', or 'libhdf5 <version> (<path>)' for the ctypes binding."""
    if _h5py is not None:
        return "h5py"
    lib = _lib()
    return "libhdf5 %d.%d.%d (%s)" % (lib.version + (lib.path,))


def _check(status, what):
    if status < 0:
        raise OSError(f"HDF5: {what} failed")
    return status


def _dims(ds):
    lib = _lib()
    sp = _check(lib.H5Dget_space(ds), "H5Dget_space")
    try:
        nd = _check(lib.H5Sget_simple_extent_ndims(sp), "H5Sget_simple_extent_ndims")
        dims = (_hsize * max(nd, 1))()
        if nd > 0:
            _check(lib.H5Sget_simple_extent_dims(sp, dims, None), "H5Sget_simple_extent_dims")
        return tuple(int(dims[i]) for i in range(nd))
    finally:
        lib.H5Sclose(sp)


def _dtype(ds):
    lib = _lib()
    ty = _check(lib.H5Dget_type(ds), "H5Dget_type")
    try:
        cls, size = lib.H5Tget_class(ty), int(lib.H5Tget_size(ty))
        if cls == _T_FLOAT and size in (4, 8):
            return np.dtype(f"f{size}")
        if cls == _T_INTEGER and size in (1, 2, 4, 8):
            kind = "u" if lib.H5Tget_sign(ty) == _T_SGN_NONE else "i"
            return np.dtype(f"{kind}{size}")
        raise TypeError(f"HDF5 dataset of type class {cls}, size {size}: only integer and "
                        f"floating-point datasets are supported without h5py")
    finally:
        lib.H5Tclose(ty)


class Dataset:
    """A dataset of an open file: ``shape``, ``dtype``, ``[...]`` and ``np.array(d)``."""

    def __init__(self, file, name):
        self.file, self.name = file, name
        lib = _lib()
        ds = lib.H5Dopen2(file._fid, name.encode(), _P_DEFAULT)
        if ds < 0:
            raise KeyError(f"Unable to open object (object '{name}' doesn't exist)")
        try:
            self.shape = _dims(ds)
            self.dtype = _dtype(ds)
        finally:
            lib.H5Dclose(ds)

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    def __len__(self):
        if not self.shape:
            raise TypeError("Attempt to take len() of scalar dataset")
        return self.shape[0]

    def read(self):
        lib = _lib()
        out = np.empty(self.shape, dtype=self.dtype)
        if out.size == 0:
            return out
        ds = _check(lib.H5Dopen2(self.file._fid, self.name.encode(), _P_DEFAULT), "H5Dopen2")
        try:
            _check(lib.H5Dread(ds, lib.types[self.dtype], _S_ALL, _S_ALL, _P_DEFAULT,
                               out.ctypes.data), f"reading '{self.name}'")
        finally:
            lib.H5Dclose(ds)
        return out

    def __getitem__(self, key):
        return self.read()[key]

    def __array__(self, dtype=None, copy=None):
        a = self.read()
        return a if dtype is None else a.astype(dtype)


class _File:
    """h5py.File's subset over the HDF5 C library (flat files of numeric datasets)."""

    def __init__(self, name, mode="r"):
        lib = _lib()
        self.filename = os.fspath(name)
        path = self.filename.encode()
        if mode == "w":
            fid = lib.H5Fcreate(path, _ACC_TRUNC, _P_DEFAULT, _P_DEFAULT)
        elif mode in ("r", "r+", "a"):
            if mode == "a" and not os.path.exists(self.filename):
                fid = lib.H5Fcreate(path, _ACC_TRUNC, _P_DEFAULT, _P_DEFAULT)
            else:
                fid = lib.H5Fopen(path, _ACC_RDONLY if mode == "r" else _ACC_RDWR, _P_DEFAULT)
        else:
            raise ValueError(f"Invalid mode; must be one of r, r+, w, a (got {mode!r})")
        if fid < 0:
            what = "create" if mode == "w" else "open"
            raise OSError(f"Unable to {what} file (name = '{self.filename}')")
        self._fid, self.mode = fid, mode

    # -- context manager / lifetime
    def close(self):
        if self._fid is not None and self._fid >= 0:
            _lib().H5Fclose(self._fid)
        self._fid = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _open(self):
        if self._fid is None:
            raise ValueError("Invalid file (the file is closed)")
        return self._fid

    # -- names
    def keys(self):
        lib = _lib()
        fid = self._open()
        info = (ctypes.c_uint64 * 8)()  # H5G_info_t: storage type, nlinks, max_corder, mounted
        _check(lib.H5Gget_info(fid, info), "H5Gget_info")
        nlinks = int(info[1])
        names = []
        for i in range(nlinks):
            n = _check(lib.H5Lget_name_by_idx(fid, b".", _INDEX_NAME, _ITER_INC, i, None, 0,
                                              _P_DEFAULT), "H5Lget_name_by_idx")
            buf = ctypes.create_string_buffer(n + 1)
            _check(lib.H5Lget_name_by_idx(fid, b".", _INDEX_NAME, _ITER_INC, i, buf, n + 1,
                                          _P_DEFAULT), "H5Lget_name_by_idx")
            names.append(buf.value.decode())
        return names

    def __iter__(self):
        return iter(self.keys())

    def __len__(self):
        return len(self.keys())

    def __contains__(self, name):
        lib = _lib()
        return lib.H5Lexists(self._open(), str(name).encode(), _P_DEFAULT) > 0

    # -- datasets
    def __getitem__(self, name):
        self._open()
        return Dataset(self, str(name))

    def create_dataset(self, name, data=None, dtype=None):
        if data is None:
            raise TypeError("create_dataset needs data= without h5py")
        self._write(name, np.asarray(data, dtype=dtype))
        return Dataset(self, name)

    def __setitem__(self, name, data):
        self._write(name, np.asarray(data))

    def _write(self, name, a):
        if a.dtype == np.bool_:
            a = a.astype(np.uint8)
        if a.dtype.byteorder not in ("=", "|"):
            a = a.astype(a.dtype.newbyteorder("="))
        a = np.require(a, requirements="C")  # keeps 0-d arrays 0-d (ascontiguousarray does not)
        lib = _lib()
        if a.dtype not in lib.types:
            raise TypeError(f"cannot store dtype {a.dtype} without h5py (integer and "
                            f"floating-point arrays only)")
        fid = self._open()
        if lib.H5Lexists(fid, name.encode(), _P_DEFAULT) > 0:
            raise ValueError(f"Unable to create dataset (name already exists): '{name}'")
        if a.ndim == 0:
            sp = _check(lib.H5Screate(_S_SCALAR), "H5Screate")
        else:
            dims = (_hsize * a.ndim)(*a.shape)
            sp = _check(lib.H5Screate_simple(a.ndim, dims, None), "H5Screate_simple")
        try:
            ty = lib.types[a.dtype]
            ds = _check(lib.H5Dcreate2(fid, name.encode(), ty, sp, _P_DEFAULT, _P_DEFAULT,
                                       _P_DEFAULT), f"creating dataset '{name}'")
            try:
                if a.size:
                    _check(lib.H5Dwrite(ds, ty, _S_ALL, _S_ALL, _P_DEFAULT, a.ctypes.data),
                           f"writing dataset '{name}'")
            finally:
                lib.H5Dclose(ds)
        finally:
            lib.H5Sclose(sp)

    def __delitem__(self, name):
        _check(_lib().H5Ldelete(self._open(), name.encode(), _P_DEFAULT), f"deleting '{name}'")


def File(name, mode="r"):
    """h5py.File when h5py is importable, else the ctypes binding's file object."""
    if _h5py is not None:
        return _h5py.File(name, mode)
    return _File(name, mode)
