"""HDF5 files for CarTens.store / CarTens.read (richmol/field.py:1304-1848).

The reference writes and reads its tensors with h5py; this package binds the
HDF5 C library itself (1.10 API, through ctypes), so files move both ways
between the two: groups, datasets and attributes in h5py's conventions --
complex numbers as the compound {r, i}, Python strings as variable-length
UTF-8 strings, numeric lists as arrays.  The library is looked up as
``$RICHMOL_AMD_HDF5``, then ``libhdf5`` on the loader path, then the conda
copy under /opt/conda/lib.

``json_dumps`` / ``json_loads`` are the attribute encoding of
richmol/json_ext.py:55-156: dictionary keys written as strings (floats
rounded to one decimal) and read back as int, then float, then string;
complex numbers as {"real", "imag", "__complex__"}; arrays as {"val",
"__ndarray__"} (read back as lists).
"""
import ctypes as C
import ctypes.util
import json
import os
from collections.abc import Mapping

import numpy as np

_hid = C.c_int64
_hsize = C.c_ulonglong

# values of the HDF5 1.10 headers
P_DEFAULT = 0
S_ALL = 0
F_ACC_RDONLY, F_ACC_RDWR, F_ACC_TRUNC, F_ACC_EXCL = 0, 1, 2, 4
I_GROUP, I_DATASET = 2, 5
T_INTEGER, T_FLOAT, T_STRING, T_COMPOUND = 0, 1, 3, 6
T_SGN_NONE = 0
T_CSET_ASCII, T_CSET_UTF8 = 0, 1
T_STR_NULLTERM, T_STR_NULLPAD = 0, 1
T_VARIABLE = C.c_size_t(-1).value
S_SCALAR = 0
INDEX_NAME, ITER_INC = 0, 0

_LIB = None
_ITER = C.CFUNCTYPE(C.c_int, _hid, C.c_char_p, C.c_void_p, C.c_void_p)


class HDF5Error(RuntimeError):
    pass


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


def lib():
    """The HDF5 library, loaded and initialised on first use."""
    global _LIB
    if _LIB is not None:
        return _LIB
    names = [os.environ.get("RICHMOL_AMD_HDF5"), ctypes.util.find_library("hdf5"), "libhdf5.so",
             "/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.103"]
    L, err = None, []
    for n in names:
        if not n:
            continue
        try:
            L = C.CDLL(n)
            break
        except OSError as e:
            err.append(f"{n}: {e}")
    if L is None:
        raise ImportError("the HDF5 C library was not found (set RICHMOL_AMD_HDF5 to libhdf5.so): "
                          + "; ".join(err))
    H, U, I, Z = _hid, C.c_uint, C.c_int, C.c_size_t
    pc, vp = C.c_char_p, C.c_void_p
    ph = C.POINTER(_hsize)
    for n, r, a in [
            ("H5open", I, []), ("H5get_libversion", I, [C.POINTER(U)] * 3),
            ("H5Eset_auto2", I, [H, vp, vp]),
            ("H5Fopen", H, [pc, U, H]), ("H5Fcreate", H, [pc, U, H, H]), ("H5Fclose", I, [H]),
            ("H5Gopen2", H, [H, pc, H]), ("H5Gcreate2", H, [H, pc, H, H, H]), ("H5Gclose", I, [H]),
            ("H5Oopen", H, [H, pc, H]), ("H5Oclose", I, [H]), ("H5Iget_type", I, [H]),
            ("H5Lexists", I, [H, pc, H]), ("H5Ldelete", I, [H, pc, H]),
            ("H5Literate", I, [H, I, I, ph, _ITER, vp]),
            ("H5Aiterate2", I, [H, I, I, ph, _ITER, vp]),
            ("H5Aexists", I, [H, pc]), ("H5Aopen", H, [H, pc, H]), ("H5Acreate2", H, [H, pc, H, H, H, H]),
            ("H5Aread", I, [H, H, vp]), ("H5Awrite", I, [H, H, vp]), ("H5Adelete", I, [H, pc]),
            ("H5Aget_type", H, [H]), ("H5Aget_space", H, [H]), ("H5Aclose", I, [H]),
            ("H5Dopen2", H, [H, pc, H]), ("H5Dcreate2", H, [H, pc, H, H, H, H, H]),
            ("H5Dread", I, [H, H, H, H, H, vp]), ("H5Dwrite", I, [H, H, H, H, H, vp]),
            ("H5Dget_type", H, [H]), ("H5Dget_space", H, [H]), ("H5Dclose", I, [H]),
            ("H5Dvlen_reclaim", I, [H, H, H, vp]),
            ("H5Screate", H, [I]), ("H5Screate_simple", H, [I, ph, ph]), ("H5Sclose", I, [H]),
            ("H5Sget_simple_extent_ndims", I, [H]), ("H5Sget_simple_extent_dims", I, [H, ph, ph]),
            ("H5Tcopy", H, [H]), ("H5Tcreate", H, [I, Z]), ("H5Tinsert", I, [H, pc, Z, H]),
            ("H5Tclose", I, [H]), ("H5Tget_class", I, [H]), ("H5Tget_size", Z, [H]),
            ("H5Tset_size", I, [H, Z]), ("H5Tget_sign", I, [H]), ("H5Tis_variable_str", I, [H]),
            ("H5Tget_cset", I, [H]), ("H5Tset_cset", I, [H, I]), ("H5Tset_strpad", I, [H, I]),
            ("H5Tget_nmembers", I, [H]), ("H5Tget_member_name", vp, [H, U]),
            ("H5Tget_member_type", H, [H, U]), ("H5Tget_member_offset", Z, [H, U]),
            ("H5free_memory", I, [vp])]:
        _sig(L, n, r, *a)
    if L.H5open() < 0:
        raise ImportError("H5open failed")
    L.H5Eset_auto2(0, None, None)   # errors come back as return codes, not printed
    maj, mnr, rel = C.c_uint(), C.c_uint(), C.c_uint()
    L.H5get_libversion(C.byref(maj), C.byref(mnr), C.byref(rel))
    if (maj.value, mnr.value) < (1, 10):
        raise ImportError(f"HDF5 {maj.value}.{mnr.value}.{rel.value}: 1.10 or later is needed")
    L._native = {k: _hid.in_dll(L, k + "_g").value for k in (
        "H5T_NATIVE_DOUBLE", "H5T_NATIVE_FLOAT", "H5T_NATIVE_INT8", "H5T_NATIVE_INT16", "H5T_NATIVE_INT32",
        "H5T_NATIVE_INT64", "H5T_NATIVE_UINT8", "H5T_NATIVE_UINT16", "H5T_NATIVE_UINT32",
        "H5T_NATIVE_UINT64", "H5T_C_S1")}
    _LIB = L
    return L


def _ck(rc, what):
    if rc < 0:
        raise HDF5Error(f"HDF5: {what} failed")
    return rc


def _int_type(dt):
    L = lib()
    sign = "UINT" if dt.kind == "u" else "INT"
    return L._native[f"H5T_NATIVE_{sign}{8 * dt.itemsize}"]


class _Type:
    """A memory/file datatype for a numpy dtype, closed on exit."""

    def __init__(self, dt, vlen_str=False):
        L = lib()
        self.own = []
        if vlen_str:
            t = L.H5Tcopy(L._native["H5T_C_S1"])
            L.H5Tset_size(t, T_VARIABLE)
            L.H5Tset_cset(t, T_CSET_UTF8)
            self.own.append(t)
        elif dt.kind == "c":
            f = L._native["H5T_NATIVE_DOUBLE" if dt.itemsize == 16 else "H5T_NATIVE_FLOAT"]
            t = L.H5Tcreate(T_COMPOUND, dt.itemsize)
            _ck(L.H5Tinsert(t, b"r", 0, f), "H5Tinsert")
            _ck(L.H5Tinsert(t, b"i", dt.itemsize // 2, f), "H5Tinsert")
            self.own.append(t)
        elif dt.kind == "f":
            t = L._native["H5T_NATIVE_DOUBLE" if dt.itemsize == 8 else "H5T_NATIVE_FLOAT"]
        elif dt.kind in "iu":
            t = _int_type(dt)
        elif dt.kind == "b":
            t = L._native["H5T_NATIVE_INT8"]
        elif dt.kind == "S":
            t = L.H5Tcopy(L._native["H5T_C_S1"])
            L.H5Tset_size(t, max(dt.itemsize, 1))
            L.H5Tset_strpad(t, T_STR_NULLPAD)
            self.own.append(t)
        else:
            raise TypeError(f"no HDF5 type for {dt}")
        self.id = t

    def close(self):
        for t in self.own:
            lib().H5Tclose(t)
        self.own = []


def _space_shape(space):
    L = lib()
    nd = L.H5Sget_simple_extent_ndims(space)
    if nd <= 0:
        return ()
    dims = (_hsize * nd)()
    L.H5Sget_simple_extent_dims(space, dims, None)
    return tuple(int(d) for d in dims)


def _read(obj, is_attr):
    """Contents of a dataset or attribute as numpy (strings as str)."""
    L = lib()
    ftype = (L.H5Aget_type if is_attr else L.H5Dget_type)(obj)
    space = (L.H5Aget_space if is_attr else L.H5Dget_space)(obj)
    reader = (lambda mt, buf: L.H5Aread(obj, mt, buf)) if is_attr else \
        (lambda mt, buf: L.H5Dread(obj, mt, S_ALL, S_ALL, P_DEFAULT, buf))
    try:
        shape = _space_shape(space)
        cls, size = L.H5Tget_class(ftype), L.H5Tget_size(ftype)
        if cls == T_STRING and L.H5Tis_variable_str(ftype) > 0:
            n = int(np.prod(shape)) if shape else 1
            mt = L.H5Tcopy(L._native["H5T_C_S1"])
            L.H5Tset_size(mt, T_VARIABLE)
            L.H5Tset_cset(mt, L.H5Tget_cset(ftype))
            buf = (C.c_char_p * n)()
            try:
                _ck(reader(mt, buf), "read")
                out = [(b or b"").decode("utf-8") for b in buf]
                L.H5Dvlen_reclaim(mt, space, P_DEFAULT, buf)
            finally:
                L.H5Tclose(mt)
            if not shape:
                return out[0]
            return np.array(out, dtype=object).reshape(shape)
        if cls == T_STRING:
            dt = np.dtype(f"S{size}")
        elif cls == T_FLOAT:
            dt = np.dtype("f8" if size == 8 else "f4")
        elif cls == T_INTEGER:
            dt = np.dtype(("u" if L.H5Tget_sign(ftype) == T_SGN_NONE else "i") + str(size))
        elif cls == T_COMPOUND:
            names = []
            for k in range(L.H5Tget_nmembers(ftype)):
                p = L.H5Tget_member_name(ftype, k)
                names.append(C.string_at(p).decode())
                L.H5free_memory(p)
            if names != ["r", "i"]:
                raise HDF5Error(f"compound type with members {names} is not a complex number")
            mtp = L.H5Tget_member_type(ftype, 0)
            dt = np.dtype("c16" if L.H5Tget_size(mtp) == 8 else "c8")
            L.H5Tclose(mtp)
        else:
            raise HDF5Error(f"HDF5 type class {cls} is not read by this module")
        arr = np.empty(shape, dtype=dt)
        mt = _Type(dt)
        try:
            _ck(reader(mt.id, arr.ctypes.data_as(C.c_void_p)), "read")
        finally:
            mt.close()
        return arr[()] if not shape else arr
    finally:
        L.H5Tclose(ftype)
        L.H5Sclose(space)


def _as_array(val):
    """(array, is vlen string) for a value h5py would store natively, or
    None when h5py would refuse it (TypeError: the caller then writes JSON)."""
    if isinstance(val, (dict, Mapping)) or val is None:
        return None
    if isinstance(val, str):
        return np.array(val, dtype=object), True
    try:
        arr = np.asarray(val)
    except (ValueError, TypeError):
        return None
    if arr.dtype.kind == "U":
        return arr.astype(object), True
    if arr.dtype.kind == "O":
        if arr.size and all(isinstance(x, str) for x in arr.ravel()):
            return arr, True
        return None
    if arr.dtype.kind in "biufcS":
        return arr, False
    return None


def _write(loc, name, val, is_attr):
    L = lib()
    got = _as_array(val)
    if got is None:
        raise TypeError(f"no HDF5 form for a value of type {type(val).__name__}")
    arr, vstr = got
    if vstr:
        n = arr.size
        enc = [s.encode("utf-8") for s in arr.ravel()]
        buf = (C.c_char_p * max(n, 1))(*enc)
        dt = None
    else:
        if arr.dtype.kind == "b":
            arr = arr.astype(np.int8)
        arr = np.array(arr, order="C")     # (ascontiguousarray would make a scalar 1-d)
        buf = arr.ctypes.data_as(C.c_void_p)
        dt = arr.dtype
    typ = _Type(dt, vlen_str=vstr)
    if arr.ndim == 0:
        space = L.H5Screate(S_SCALAR)
    else:
        dims = (_hsize * arr.ndim)(*arr.shape)
        space = L.H5Screate_simple(arr.ndim, dims, None)
    try:
        if is_attr:
            if L.H5Aexists(loc, name.encode()) > 0:
                _ck(L.H5Adelete(loc, name.encode()), f"deleting attribute '{name}'")
            a = _ck(L.H5Acreate2(loc, name.encode(), typ.id, space, P_DEFAULT, P_DEFAULT),
                    f"creating attribute '{name}'")
            try:
                _ck(L.H5Awrite(a, typ.id, buf), f"writing attribute '{name}'")
            finally:
                L.H5Aclose(a)
        else:
            d = _ck(L.H5Dcreate2(loc, name.encode(), typ.id, space, P_DEFAULT, P_DEFAULT, P_DEFAULT),
                    f"creating dataset '{name}'")
            try:
                _ck(L.H5Dwrite(d, typ.id, S_ALL, S_ALL, P_DEFAULT, buf), f"writing dataset '{name}'")
            finally:
                L.H5Dclose(d)
    finally:
        typ.close()
        L.H5Sclose(space)


class Attrs:
    """Attributes of a group or dataset (read, write, iterate)."""

    def __init__(self, oid):
        self._id = oid

    def keys(self):
        names = []
        cb = _ITER(lambda loc, name, info, op: names.append(name.decode()) or 0)
        _ck(lib().H5Aiterate2(self._id, INDEX_NAME, ITER_INC, None, cb, None), "H5Aiterate2")
        return names

    def __contains__(self, name):
        return lib().H5Aexists(self._id, name.encode()) > 0

    def __getitem__(self, name):
        L = lib()
        if name not in self:
            raise KeyError(name)
        a = _ck(L.H5Aopen(self._id, name.encode(), P_DEFAULT), f"opening attribute '{name}'")
        try:
            return _read(a, True)
        finally:
            L.H5Aclose(a)

    def __setitem__(self, name, val):
        _write(self._id, name, val, True)

    def items(self):
        return [(k, self[k]) for k in self.keys()]


class Dataset:
    def __init__(self, oid):
        self._id = oid
        self.attrs = Attrs(oid)

    def read(self):
        return _read(self._id, False)

    def __getitem__(self, idx):
        return self.read()[idx]

    def __array__(self, dtype=None, copy=None):
        a = self.read()
        return a if dtype is None else a.astype(dtype)

    def close(self):
        if self._id:
            lib().H5Oclose(self._id)
            self._id = 0


class Group:
    def __init__(self, oid, children):
        self._id = oid
        self._children = children     # opened objects, closed with the file
        self.attrs = Attrs(oid)

    def keys(self):
        names = []
        cb = _ITER(lambda loc, name, info, op: names.append(name.decode()) or 0)
        _ck(lib().H5Literate(self._id, INDEX_NAME, ITER_INC, None, cb, None), "H5Literate")
        return names

    def __contains__(self, name):
        return lib().H5Lexists(self._id, name.encode(), P_DEFAULT) > 0

    def __getitem__(self, name):
        L = lib()
        if name not in self:
            raise KeyError(name)
        o = _ck(L.H5Oopen(self._id, name.encode(), P_DEFAULT), f"opening '{name}'")
        t = L.H5Iget_type(o)
        obj = Group(o, self._children) if t == I_GROUP else Dataset(o) if t == I_DATASET else None
        if obj is None:
            L.H5Oclose(o)
            raise HDF5Error(f"'{name}' is neither a group nor a dataset")
        self._children.append(obj)
        return obj

    def __delitem__(self, name):
        if name not in self:
            raise KeyError(name)
        _ck(lib().H5Ldelete(self._id, name.encode(), P_DEFAULT), f"deleting '{name}'")

    def create_group(self, name):
        L = lib()
        if name in self:
            raise ValueError(f"'{name}' exists")
        g = _ck(L.H5Gcreate2(self._id, name.encode(), P_DEFAULT, P_DEFAULT, P_DEFAULT),
                f"creating group '{name}'")
        obj = Group(g, self._children)
        self._children.append(obj)
        return obj

    def require_group(self, name):
        return self[name] if name in self else self.create_group(name)

    def create_dataset(self, name, data):
        if name in self:
            raise ValueError(f"'{name}' exists")
        _write(self._id, name, data, False)
        return self[name]

    def close(self):
        if self._id:
            lib().H5Oclose(self._id)
            self._id = 0


class File(Group):
    """An HDF5 file: mode 'r' (read), 'r+' (read and write), 'w' (create,
    truncate) or 'a' (read and write, created when absent)."""

    def __init__(self, filename, mode="r"):
        L = lib()
        fn = os.fspath(filename).encode()
        if mode == "r":
            fid = L.H5Fopen(fn, F_ACC_RDONLY, P_DEFAULT)
        elif mode == "r+":
            fid = L.H5Fopen(fn, F_ACC_RDWR, P_DEFAULT)
        elif mode == "w":
            fid = L.H5Fcreate(fn, F_ACC_TRUNC, P_DEFAULT, P_DEFAULT)
        elif mode == "a":
            fid = L.H5Fopen(fn, F_ACC_RDWR, P_DEFAULT) if os.path.exists(filename) else \
                L.H5Fcreate(fn, F_ACC_EXCL, P_DEFAULT, P_DEFAULT)
        else:
            raise ValueError(f"unknown mode '{mode}'")
        if fid < 0:
            raise OSError(f"unable to open HDF5 file '{os.fspath(filename)}' (mode '{mode}')")
        self._fid = fid
        root = _ck(L.H5Gopen2(fid, b"/", P_DEFAULT), "opening the root group")
        super().__init__(root, [])

    def close(self):
        if self._fid:
            for o in reversed(self._children):
                o.close()
            self._children.clear()
            super().close()
            lib().H5Fclose(self._fid)
            self._fid = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def is_hdf5(filename):
    """True when the file begins with the HDF5 signature (at offset 0, 512,
    1024, ...)."""
    sig = b"\x89HDF\r\n\x1a\n"
    try:
        with open(filename, "rb") as f:
            off = 0
            while True:
                f.seek(off)
                head = f.read(8)
                if len(head) < 8:
                    return False
                if head == sig:
                    return True
                off = 512 if off == 0 else 2 * off
    except OSError:
        return False


# ---------------------------------------------------- json_ext.py:55-156
def _keys_out(node):
    out = {}
    for k, v in node.items():
        v = _keys_out(v) if isinstance(v, Mapping) else v
        if isinstance(k, (float, np.floating)):
            k = str(round(float(k), 1))
        else:
            try:
                k = int(k)
            except (ValueError, TypeError):
                pass
        out[k] = v
    return out


def _keys_in(node):
    out = {}
    for k, v in node.items():
        v = _keys_in(v) if isinstance(v, Mapping) else v
        try:
            k = int(k) if k.find(".") == -1 else round(float(k), 1)
        except ValueError:
            pass
        out[k] = v
    return out


class _Enc(json.JSONEncoder):
    def default(self, o):
        if isinstance(o, (complex, np.complexfloating)):
            return {"real": float(o.real), "imag": float(o.imag), "__complex__": True}
        if isinstance(o, np.ndarray):
            res = {"val": o.tolist(), "__ndarray__": True}
            if o.dtype.fields is not None:
                res["dtype"] = [(n, str(f[0].subdtype[0]), f[0].subdtype[1]) if f[0].subdtype is not None
                                else (n, str(f[0])) for n, f in o.dtype.fields.items()]
            return res
        if isinstance(o, np.integer):
            return int(o)
        if isinstance(o, np.floating):
            return float(o)
        if callable(o):
            return None
        return super().default(o)


def _hook(d):
    if "__complex__" in d:
        return complex(d["real"], d["imag"])
    if "__ndarray__" in d:
        if "dtype" in d:
            dt = [tuple(e) for e in d["dtype"]]
            return np.array([tuple(e) for e in d["val"]], dtype=dt)
        return d["val"]
    return d


def json_dumps(val):
    return json.dumps(_keys_out(val) if isinstance(val, Mapping) else val, cls=_Enc, sort_keys=True)


def json_loads(s):
    v = json.loads(s, object_hook=_hook)
    return _keys_in(v) if isinstance(v, Mapping) else v
