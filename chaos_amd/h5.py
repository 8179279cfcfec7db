"""A small ctypes binding of the HDF5 C library (1.10 API) for the run store: files, groups, contiguous datasets and
attributes, written with the type conventions of HDF5.jl 0.17, the reference's writer (Project.toml: HDF5 = "0.17"):

- Int64 and Float64 as the native little-endian 64-bit types; UInt32 as the native unsigned 32-bit type;
- Bool as the 8-bit bitfield H5T_NATIVE_B8;
- a String as a fixed-length UTF-8 string of its byte length (at least 1) on a scalar dataspace;
- a vector of strings as variable-length UTF-8 strings on a one-dimensional dataspace;
- an array with its dimensions reversed: a Julia (n, k) array is the file extent (k, n) holding the same bytes, so
  a numpy array is written as its transpose in C order and read back transposed (HDF5.jl stores column-major).

The library is looked up at `CL_HDF5_LIB`, then by the linker's name search, then in the conda prefix of the image
(`/opt/conda/lib`); `load()` raises when none loads. Every failed call raises `H5Error` (the library's own error
printing is switched off).
"""
import ctypes
import ctypes.util
import os

import numpy as np

hid_t = ctypes.c_int64
herr_t = ctypes.c_int
htri_t = ctypes.c_int
hsize_t = ctypes.c_uint64

H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC = 0x0000, 0x0001, 0x0002
H5F_SCOPE_LOCAL = 0
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR, H5S_SIMPLE, H5S_NULL = 0, 1, 2
H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_BITFIELD, H5T_ENUM = 0, 1, 3, 4, 8
H5T_SGN_NONE = 0
H5T_VARIABLE = ctypes.c_size_t(-1).value
H5T_CSET_UTF8 = 1
H5I_GROUP, H5I_DATASET = 2, 5
H5_INDEX_NAME, H5_ITER_INC = 0, 0

_SEARCH = ("/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.103", "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so")


class H5Error(RuntimeError):
    pass


class _H5G_info(ctypes.Structure):
    _fields_ = [("storage_type", ctypes.c_int), ("nlinks", hsize_t), ("max_corder", ctypes.c_int64),
                ("mounted", ctypes.c_bool)]


_lib = None
_T = {}


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


def load():
    """The HDF5 library, loaded once; raises H5Error naming the places searched when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    cands = [os.environ.get("CL_HDF5_LIB"), ctypes.util.find_library("hdf5"), *_SEARCH]
    err = []
    for c in cands:
        if not c:
            continue
        try:
            lib = ctypes.CDLL(c)
            break
        except OSError as e:
            err.append(f"{c}: {e}")
    else:
        raise H5Error("no HDF5 library (set CL_HDF5_LIB): " + "; ".join(err or ["nothing found"]))
    P, S, C, V = ctypes.POINTER, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p
    _sig(lib, "H5open", herr_t)
    _sig(lib, "H5get_libversion", herr_t, P(ctypes.c_uint), P(ctypes.c_uint), P(ctypes.c_uint))
    _sig(lib, "H5Eset_auto2", herr_t, hid_t, V, V)
    _sig(lib, "H5Fcreate", hid_t, S, ctypes.c_uint, hid_t, hid_t)
    _sig(lib, "H5Fopen", hid_t, S, ctypes.c_uint, hid_t)
    _sig(lib, "H5Fflush", herr_t, hid_t, ctypes.c_int)
    _sig(lib, "H5Fclose", herr_t, hid_t)
    _sig(lib, "H5Gcreate2", hid_t, hid_t, S, hid_t, hid_t, hid_t)
    _sig(lib, "H5Gopen2", hid_t, hid_t, S, hid_t)
    _sig(lib, "H5Gget_info", herr_t, hid_t, P(_H5G_info))
    _sig(lib, "H5Gclose", herr_t, hid_t)
    _sig(lib, "H5Lexists", htri_t, hid_t, S, hid_t)
    _sig(lib, "H5Lget_name_by_idx", ctypes.c_ssize_t, hid_t, S, ctypes.c_int, ctypes.c_int, hsize_t, V, C, hid_t)
    _sig(lib, "H5Oopen", hid_t, hid_t, S, hid_t)
    _sig(lib, "H5Oclose", herr_t, hid_t)
    _sig(lib, "H5Iget_type", ctypes.c_int, hid_t)
    _sig(lib, "H5Pcreate", hid_t, hid_t)
    _sig(lib, "H5Pset_create_intermediate_group", herr_t, hid_t, ctypes.c_uint)
    _sig(lib, "H5Pclose", herr_t, hid_t)
    _sig(lib, "H5Screate", hid_t, ctypes.c_int)
    _sig(lib, "H5Screate_simple", hid_t, ctypes.c_int, P(hsize_t), P(hsize_t))
    _sig(lib, "H5Sget_simple_extent_type", ctypes.c_int, hid_t)
    _sig(lib, "H5Sget_simple_extent_ndims", ctypes.c_int, hid_t)
    _sig(lib, "H5Sget_simple_extent_dims", ctypes.c_int, hid_t, P(hsize_t), P(hsize_t))
    _sig(lib, "H5Sclose", herr_t, hid_t)
    _sig(lib, "H5Tcopy", hid_t, hid_t)
    _sig(lib, "H5Tset_size", herr_t, hid_t, C)
    _sig(lib, "H5Tset_cset", herr_t, hid_t, ctypes.c_int)
    _sig(lib, "H5Tget_class", ctypes.c_int, hid_t)
    _sig(lib, "H5Tget_size", C, hid_t)
    _sig(lib, "H5Tget_sign", ctypes.c_int, hid_t)
    _sig(lib, "H5Tis_variable_str", htri_t, hid_t)
    _sig(lib, "H5Tget_super", hid_t, hid_t)
    _sig(lib, "H5Tclose", herr_t, hid_t)
    _sig(lib, "H5Dcreate2", hid_t, hid_t, S, hid_t, hid_t, hid_t, hid_t, hid_t)
    _sig(lib, "H5Dopen2", hid_t, hid_t, S, hid_t)
    _sig(lib, "H5Dwrite", herr_t, hid_t, hid_t, hid_t, hid_t, hid_t, V)
    _sig(lib, "H5Dread", herr_t, hid_t, hid_t, hid_t, hid_t, hid_t, V)
    _sig(lib, "H5Dget_type", hid_t, hid_t)
    _sig(lib, "H5Dget_space", hid_t, hid_t)
    _sig(lib, "H5Dvlen_reclaim", herr_t, hid_t, hid_t, hid_t, V)
    _sig(lib, "H5Dclose", herr_t, hid_t)
    _sig(lib, "H5Acreate2", hid_t, hid_t, S, hid_t, hid_t, hid_t, hid_t)
    _sig(lib, "H5Aopen", hid_t, hid_t, S, hid_t)
    _sig(lib, "H5Aopen_by_idx", hid_t, hid_t, S, ctypes.c_int, ctypes.c_int, hsize_t, hid_t, hid_t)
    _sig(lib, "H5Aget_num_attrs", ctypes.c_int, hid_t)
    _sig(lib, "H5Aget_name", ctypes.c_ssize_t, hid_t, C, V)
    _sig(lib, "H5Awrite", herr_t, hid_t, hid_t, V)
    _sig(lib, "H5Aread", herr_t, hid_t, hid_t, V)
    _sig(lib, "H5Aget_type", hid_t, hid_t)
    _sig(lib, "H5Aget_space", hid_t, hid_t)
    _sig(lib, "H5Aexists", htri_t, hid_t, S)
    _sig(lib, "H5Adelete", herr_t, hid_t, S)
    _sig(lib, "H5Aclose", herr_t, hid_t)
    if lib.H5open() < 0:
        raise H5Error("H5open failed")
    lib.H5Eset_auto2(0, None, None)
    for nm in ("H5T_NATIVE_DOUBLE", "H5T_NATIVE_INT64", "H5T_NATIVE_UINT32", "H5T_NATIVE_INT32", "H5T_NATIVE_UINT64",
               "H5T_NATIVE_UINT8", "H5T_NATIVE_INT8", "H5T_NATIVE_B8", "H5T_NATIVE_FLOAT", "H5T_C_S1",
               "H5P_CLS_LINK_CREATE_ID"):
        _T[nm] = hid_t.in_dll(lib, nm + "_g").value
    _lib = lib
    return lib


def version():
    lib = load()
    a, b, c = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
    lib.H5get_libversion(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def _ck(v, what):
    if v < 0:
        raise H5Error(f"HDF5: {what} failed")
    return v


_NP = {np.dtype(np.float64): "H5T_NATIVE_DOUBLE", np.dtype(np.int64): "H5T_NATIVE_INT64",
       np.dtype(np.uint32): "H5T_NATIVE_UINT32", np.dtype(np.int32): "H5T_NATIVE_INT32",
       np.dtype(np.uint64): "H5T_NATIVE_UINT64", np.dtype(np.uint8): "H5T_NATIVE_UINT8",
       np.dtype(np.int8): "H5T_NATIVE_INT8", np.dtype(np.float32): "H5T_NATIVE_FLOAT"}
JULIA_NAME = {np.dtype(np.float64): "Float64", np.dtype(np.int64): "Int64", np.dtype(np.uint32): "UInt32",
              np.dtype(np.int32): "Int32", np.dtype(np.uint64): "UInt64", np.dtype(np.uint8): "UInt8",
              np.dtype(np.int8): "Int8", np.dtype(np.float32): "Float32"}


def _space(file_dims):
    lib = _lib
    if file_dims is None:
        return _ck(lib.H5Screate(H5S_SCALAR), "H5Screate")
    d = (hsize_t * max(1, len(file_dims)))(*file_dims)
    return _ck(lib.H5Screate_simple(len(file_dims), d, d), "H5Screate_simple")


def _read_space(sp):
    """None for a scalar dataspace, () for a null one, else the file extents."""
    lib = _lib
    cls = lib.H5Sget_simple_extent_type(sp)
    if cls == H5S_SCALAR:
        return None
    if cls == H5S_NULL:
        return ()
    nd = _ck(lib.H5Sget_simple_extent_ndims(sp), "H5Sget_simple_extent_ndims")
    d = (hsize_t * max(1, nd))()
    lib.H5Sget_simple_extent_dims(sp, d, None)
    return tuple(int(d[i]) for i in range(nd))


def _np_of_type(t):
    """The numpy dtype for a numeric HDF5 type, "bool" for an 8-bit bitfield or enum, "str" / "vstr" for strings."""
    lib = _lib
    cls, size = lib.H5Tget_class(t), lib.H5Tget_size(t)
    if cls == H5T_FLOAT:
        return np.dtype({8: np.float64, 4: np.float32}[size])
    if cls == H5T_INTEGER:
        unsigned = lib.H5Tget_sign(t) == H5T_SGN_NONE
        return np.dtype(("u" if unsigned else "i") + str(size))
    if cls in (H5T_BITFIELD, H5T_ENUM) and size == 1:
        return "bool"
    if cls == H5T_STRING:
        return "vstr" if lib.H5Tis_variable_str(t) > 0 else "str"
    raise H5Error(f"unsupported HDF5 type class {cls} (size {size})")


def _encode_value(v):
    """(memory type, owned type to close or 0, file dims or None, buffer, keep-alive) of an attribute value."""
    if isinstance(v, (bool, np.bool_)):
        buf = (ctypes.c_uint8 * 1)(1 if v else 0)
        return _T["H5T_NATIVE_B8"], 0, None, buf, None
    if isinstance(v, (int, np.integer)) and not isinstance(v, (bool, np.bool_)):
        buf = (ctypes.c_int64 * 1)(int(v))
        return _T["H5T_NATIVE_INT64"], 0, None, buf, None
    if isinstance(v, (float, np.floating)):
        buf = (ctypes.c_double * 1)(float(v))
        return _T["H5T_NATIVE_DOUBLE"], 0, None, buf, None
    if isinstance(v, str):
        b = v.encode("utf-8")
        t = _ck(_lib.H5Tcopy(_T["H5T_C_S1"]), "H5Tcopy")
        _lib.H5Tset_size(t, max(len(b), 1))
        _lib.H5Tset_cset(t, H5T_CSET_UTF8)
        buf = ctypes.create_string_buffer(b, max(len(b), 1))
        return t, t, None, buf, None
    if isinstance(v, (list, tuple)) and all(isinstance(s, str) for s in v) and (len(v) > 0 or isinstance(v, list)):
        enc = [s.encode("utf-8") for s in v]
        t = _ck(_lib.H5Tcopy(_T["H5T_C_S1"]), "H5Tcopy")
        _lib.H5Tset_size(t, H5T_VARIABLE)
        _lib.H5Tset_cset(t, H5T_CSET_UTF8)
        arr = (ctypes.c_char_p * max(1, len(enc)))(*enc)
        return t, t, (len(enc),), arr, enc
    a = np.asarray(v)
    if a.dtype not in _NP:
        raise H5Error(f"attribute value of type {type(v).__name__} is not written natively")
    a = np.ascontiguousarray(a.T)
    return _T[_NP[a.dtype]], 0, a.shape, a.ctypes.data_as(ctypes.c_void_p), a


def _decode(obj_id, read, get_type, get_space, what):
    """Read an attribute or dataset as Python / numpy values (Julia shapes for arrays)."""
    lib = _lib
    t = _ck(get_type(obj_id), what + " type")
    sp = _ck(get_space(obj_id), what + " space")
    try:
        kind = _np_of_type(t)
        dims = _read_space(sp)
        n = 1 if dims is None else int(np.prod(dims)) if dims else 0
        if kind == "str":
            size = lib.H5Tget_size(t)
            buf = ctypes.create_string_buffer(size * max(n, 1) + 1)
            if n:
                _ck(read(obj_id, t, buf), what + " read")
            raw = buf.raw[:size * max(n, 1)]
            vals = [raw[i * size:(i + 1) * size].split(b"\0", 1)[0].decode("utf-8") for i in range(max(n, 1))]
            return vals[0] if dims is None else ("" if dims == () else vals[:n])
        if kind == "vstr":
            arr = (ctypes.c_char_p * max(n, 1))()
            if n:
                _ck(read(obj_id, t, arr), what + " read")
            vals = [(arr[i] or b"").decode("utf-8") for i in range(n)]
            if n:
                lib.H5Dvlen_reclaim(t, sp, H5P_DEFAULT, arr)
            return vals[0] if dims is None else vals
        if kind == "bool":
            buf = (ctypes.c_uint8 * max(n, 1))()
            if n:  # read in the stored one-byte type: HDF5 converts no bitfield to an integer
                _ck(read(obj_id, t, buf), what + " read")
            vals = [bool(buf[i]) for i in range(n)]
            return vals[0] if dims is None else np.array(vals, dtype=bool).reshape(dims).T
        out = np.empty(dims if dims else (0,) if dims == () else (), dtype=kind)
        if n:
            _ck(read(obj_id, _T[_NP[kind]], out.ctypes.data_as(ctypes.c_void_p)), what + " read")
        if dims is None:
            return out.item() if kind.kind == "f" else int(out.item())
        return out.T
    finally:
        lib.H5Sclose(sp)
        lib.H5Tclose(t)


class Node:
    """A file or group: children by name, attributes, datasets."""

    def __init__(self, hid, close):
        self.id, self._close = hid, close

    def _lcpl(self):
        p = _ck(_lib.H5Pcreate(_T["H5P_CLS_LINK_CREATE_ID"]), "H5Pcreate")
        _lib.H5Pset_create_intermediate_group(p, 1)
        return p

    def create_group(self, path):
        p = self._lcpl()
        try:
            g = _ck(_lib.H5Gcreate2(self.id, path.encode(), p, H5P_DEFAULT, H5P_DEFAULT), f"create group {path}")
        finally:
            _lib.H5Pclose(p)
        return Node(g, _lib.H5Gclose)

    def group(self, path):
        return Node(_ck(_lib.H5Gopen2(self.id, path.encode(), H5P_DEFAULT), f"open group {path}"), _lib.H5Gclose)

    def has(self, name):
        return _lib.H5Lexists(self.id, name.encode(), H5P_DEFAULT) > 0

    def is_dataset(self, name):
        o = _ck(_lib.H5Oopen(self.id, name.encode(), H5P_DEFAULT), f"open {name}")
        try:
            return _lib.H5Iget_type(o) == H5I_DATASET
        finally:
            _lib.H5Oclose(o)

    def keys(self):
        """Child names in HDF5's name order (the order HDF5.jl's keys lists them in)."""
        info = _H5G_info()
        _ck(_lib.H5Gget_info(self.id, ctypes.byref(info)), "H5Gget_info")
        out = []
        for i in range(info.nlinks):
            n = _lib.H5Lget_name_by_idx(self.id, b".", H5_INDEX_NAME, H5_ITER_INC, i, None, 0, H5P_DEFAULT)
            buf = ctypes.create_string_buffer(_ck(n, "H5Lget_name_by_idx") + 1)
            _lib.H5Lget_name_by_idx(self.id, b".", H5_INDEX_NAME, H5_ITER_INC, i, buf, n + 1, H5P_DEFAULT)
            out.append(buf.value.decode("utf-8"))
        return out

    def write_dataset(self, path, data):
        """A contiguous dataset holding the numpy array `data` (Julia shape; stored with reversed extents)."""
        a = np.asarray(data)
        if a.dtype not in _NP:
            raise H5Error(f"dataset {path}: element type {a.dtype} is not supported")
        c = np.ascontiguousarray(a.T)
        sp = _space(c.shape)
        p = self._lcpl()
        try:
            d = _ck(_lib.H5Dcreate2(self.id, path.encode(), _T[_NP[a.dtype]], sp, p, H5P_DEFAULT, H5P_DEFAULT),
                    f"create dataset {path}")
        finally:
            _lib.H5Pclose(p)
            _lib.H5Sclose(sp)
        try:
            if c.size:
                _ck(_lib.H5Dwrite(d, _T[_NP[a.dtype]], H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                  c.ctypes.data_as(ctypes.c_void_p)), f"write dataset {path}")
        except H5Error:
            _lib.H5Dclose(d)
            raise
        return Node(d, _lib.H5Dclose)

    def dataset(self, path):
        return Node(_ck(_lib.H5Dopen2(self.id, path.encode(), H5P_DEFAULT), f"open dataset {path}"), _lib.H5Dclose)

    def read(self):
        """The dataset's values (this node must be a dataset), in the Julia shape."""
        rd = lambda i, t, b: _lib.H5Dread(i, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, b)
        return _decode(self.id, rd, _lib.H5Dget_type, _lib.H5Dget_space, "dataset")

    def layout(self):
        """(Julia element type name, Julia size) of this dataset."""
        t = _ck(_lib.H5Dget_type(self.id), "dataset type")
        sp = _ck(_lib.H5Dget_space(self.id), "dataset space")
        try:
            kind = _np_of_type(t)
            dims = _read_space(sp)
        finally:
            _lib.H5Sclose(sp)
            _lib.H5Tclose(t)
        name = JULIA_NAME.get(kind, "Bool" if kind == "bool" else "String")
        return name, tuple(reversed(dims or ()))

    def set_attr(self, name, value):
        t, owned, dims, buf, keep = _encode_value(value)
        sp = _space(dims)
        try:
            a = _ck(_lib.H5Acreate2(self.id, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT), f"create attribute {name}")
            try:
                if dims is None or int(np.prod(dims)):
                    _ck(_lib.H5Awrite(a, t, ctypes.cast(buf, ctypes.c_void_p)), f"write attribute {name}")
            finally:
                _lib.H5Aclose(a)
        finally:
            _lib.H5Sclose(sp)
            if owned:
                _lib.H5Tclose(owned)
        del keep

    def del_attr(self, name):
        _ck(_lib.H5Adelete(self.id, name.encode()), f"delete attribute {name}")

    def attr_names(self):
        n = _ck(_lib.H5Aget_num_attrs(self.id), "H5Aget_num_attrs")
        out = []
        for i in range(n):
            a = _ck(_lib.H5Aopen_by_idx(self.id, b".", H5_INDEX_NAME, H5_ITER_INC, i, H5P_DEFAULT, H5P_DEFAULT),
                    "H5Aopen_by_idx")
            try:
                k = _ck(_lib.H5Aget_name(a, 0, None), "H5Aget_name")
                buf = ctypes.create_string_buffer(k + 1)
                _lib.H5Aget_name(a, k + 1, buf)
                out.append(buf.value.decode("utf-8"))
            finally:
                _lib.H5Aclose(a)
        return out

    def attr(self, name):
        a = _ck(_lib.H5Aopen(self.id, name.encode(), H5P_DEFAULT), f"open attribute {name}")
        try:
            return _decode(a, _lib.H5Aread, _lib.H5Aget_type, _lib.H5Aget_space, f"attribute {name}")
        finally:
            _lib.H5Aclose(a)

    def attrs(self):
        return {k: self.attr(k) for k in self.attr_names()}

    def close(self):
        if self.id is not None:
            self._close(self.id)
            self.id = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def open_file(path, mode="r"):
    """`mode` "w" (create, truncating), "r" (read) or "r+" (read and write)."""
    lib = load()
    p = os.fsencode(path)
    if mode == "w":
        f = lib.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
    else:
        f = lib.H5Fopen(p, H5F_ACC_RDWR if mode == "r+" else H5F_ACC_RDONLY, H5P_DEFAULT)
    _ck(f, f"open {path} ({mode})")
    root = lib.H5Gopen2(f, b"/", H5P_DEFAULT)
    if root < 0:
        lib.H5Fclose(f)
        raise H5Error(f"HDF5: open the root group of {path} failed")

    def close(g):
        lib.H5Gclose(g)
        _ck(lib.H5Fclose(f), f"close {path}")

    return Node(root, close)  # the root group is the location of the file's attributes and children
