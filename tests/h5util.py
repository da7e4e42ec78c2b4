"""Minimal HDF5 access for the tests, through the HDF5 1.10 C library with
ctypes (h5py is not installed): read a dataset into numpy, and write a
VASP-style vaspout.h5 (structure, charge, POTCAR text) in the layout the
reference's reader takes (reference: src/half_hdf5_bridge.c:10-74).  This is
an independent implementation of the file access, used to check the
library's reader and writer from outside."""
import ctypes as C
import os

import numpy as np

hid = C.c_int64
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    names = [os.environ.get("HALF_HDF5_LIBRARY"), "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103"]
    for n in names:
        if not n:
            continue
        try:
            lib = C.CDLL(n)
            break
        except OSError:
            continue
    else:
        return None
    sig = {
        "H5open": ([], C.c_int), "H5Eset_auto2": ([hid, C.c_void_p, C.c_void_p], C.c_int),
        "H5Fcreate": ([C.c_char_p, C.c_uint, hid, hid], hid), "H5Fopen": ([C.c_char_p, C.c_uint, hid], hid),
        "H5Fclose": ([hid], C.c_int), "H5Gcreate2": ([hid, C.c_char_p, hid, hid, hid], hid),
        "H5Gclose": ([hid], C.c_int), "H5Screate": ([C.c_int], hid),
        "H5Screate_simple": ([C.c_int, C.POINTER(C.c_ulonglong), C.c_void_p], hid), "H5Sclose": ([hid], C.c_int),
        "H5Sget_simple_extent_ndims": ([hid], C.c_int),
        "H5Sget_simple_extent_dims": ([hid, C.POINTER(C.c_ulonglong), C.c_void_p], C.c_int),
        "H5Dcreate2": ([hid, C.c_char_p, hid, hid, hid, hid, hid], hid), "H5Dopen2": ([hid, C.c_char_p, hid], hid),
        "H5Dwrite": ([hid, hid, hid, hid, hid, C.c_void_p], C.c_int),
        "H5Dread": ([hid, hid, hid, hid, hid, C.c_void_p], C.c_int), "H5Dclose": ([hid], C.c_int),
        "H5Dget_space": ([hid], hid), "H5Tcopy": ([hid], hid), "H5Tset_size": ([hid, C.c_size_t], C.c_int),
        "H5Tclose": ([hid], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes, f.restype = args, res
    lib.H5open()
    lib.H5Eset_auto2(0, None, None)
    _lib = lib
    return lib


def available():
    return load() is not None


def _type(lib, kind):
    return hid.in_dll(lib, {"f8": "H5T_NATIVE_DOUBLE_g", "f4": "H5T_NATIVE_FLOAT_g", "i4": "H5T_NATIVE_INT_g",
                            "s": "H5T_C_S1_g"}[kind]).value


def read(path, name, kind="f8", strlen=16):
    """Dataset `name` as a numpy array (kind f8, f4, i4) or a list of
    strings (kind s, fixed length strlen)."""
    lib = load()
    f = lib.H5Fopen(str(path).encode(), 0, 0)
    assert f >= 0, f"cannot open {path}"
    try:
        d = lib.H5Dopen2(f, name.encode(), 0)
        assert d >= 0, f"{path} has no {name}"
        sp = lib.H5Dget_space(d)
        r = lib.H5Sget_simple_extent_ndims(sp)
        dims = (C.c_ulonglong * max(r, 1))()
        lib.H5Sget_simple_extent_dims(sp, dims, None)
        lib.H5Sclose(sp)
        shape = tuple(int(dims[i]) for i in range(r))
        if kind == "s":
            t = lib.H5Tcopy(_type(lib, "s"))
            lib.H5Tset_size(t, strlen)
            buf = C.create_string_buffer(strlen * max(1, int(np.prod(shape))))
            assert lib.H5Dread(d, t, 0, 0, 0, buf) >= 0
            lib.H5Tclose(t)
            raw = buf.raw
            out = [raw[i * strlen:(i + 1) * strlen] for i in range(max(1, int(np.prod(shape))))]
        else:
            out = np.zeros(shape, {"f8": np.float64, "f4": np.float32, "i4": np.int32}[kind])
            assert lib.H5Dread(d, _type(lib, kind), 0, 0, 0, out.ctypes.data) >= 0
        lib.H5Dclose(d)
        return out
    finally:
        lib.H5Fclose(f)


def _put(lib, parent, name, kind, data, strlen=None, count=None):
    """count: number of fixed-length strings (None: one scalar string)."""
    if kind == "s":
        t = lib.H5Tcopy(_type(lib, "s"))
        lib.H5Tset_size(t, strlen)
    else:
        t = _type(lib, kind)
    arr = None
    if kind == "s":
        buf = data
        shape = () if count is None else (count,)
    else:
        arr = np.ascontiguousarray(data, {"f8": np.float64, "f4": np.float32, "i4": np.int32}[kind])
        shape = arr.shape
        buf = arr.ctypes.data
    if shape:
        dims = (C.c_ulonglong * len(shape))(*shape)
        sp = lib.H5Screate_simple(len(shape), dims, None)
    else:
        sp = lib.H5Screate(0)
    d = lib.H5Dcreate2(parent, name.encode(), t, sp, 0, 0, 0)
    assert d >= 0, name
    assert lib.H5Dwrite(d, t, 0, 0, 0, C.c_char_p(buf) if kind == "s" else buf) >= 0, name
    lib.H5Dclose(d)
    lib.H5Sclose(sp)
    if kind == "s":
        lib.H5Tclose(t)


def write_vaspout(path, lattice, species, counts, positions, grid, values, potcar_text=None,
                  group="structure/positions", direct=True, spin_axis=True, system="synthetic vaspout",
                  extra_types=0, system_count=None):
    """A vaspout.h5-like file: structure (fractional or Cartesian positions),
    charge (values x fastest, with or without the leading spin axis) and,
    if given, the POTCAR text at /input/potcar/content."""
    lib = load()
    f = lib.H5Fcreate(str(path).encode(), 2, 0, 0)
    assert f >= 0
    opened = []
    try:
        parts = group.split("/")
        g = f
        for p in parts:
            g = lib.H5Gcreate2(g, p.encode(), 0, 0, 0)
            opened.append(g)
        if system_count is None:
            _put(lib, g, "system", "s", system.encode(), len(system))
        else:  # malformed: an array of strings
            _put(lib, g, "system", "s", system.encode() * system_count, len(system), count=system_count)
        _put(lib, g, "lattice_vectors", "f8", np.asarray(lattice, float).reshape(3, 3))
        names = b"".join(s.encode().ljust(16) for s in list(species) + ["X"] * extra_types)
        _put(lib, g, "ion_types", "s", names, 16, count=len(species) + extra_types)
        _put(lib, g, "number_ion_types", "i4", np.asarray(counts))
        pos = np.asarray(positions, float).reshape(-1, 3)
        if not direct:
            pos = pos @ np.asarray(lattice, float).reshape(3, 3)
        _put(lib, g, "position_ions", "f8", pos)
        _put(lib, g, "direct_coordinates", "i4", np.array(1 if direct else 0))
        ch = lib.H5Gcreate2(f, b"charge", 0, 0, 0)
        opened.append(ch)
        _put(lib, ch, "grid", "i4", np.asarray(grid))
        v = np.asarray(values, float).reshape(grid[2], grid[1], grid[0])  # x fastest
        _put(lib, ch, "charge", "f8", v[None] if spin_axis else v)
        if potcar_text is not None:
            if parts[0] == "input":  # the structure group already made /input
                i = opened[0]
            else:
                i = lib.H5Gcreate2(f, b"input", 0, 0, 0)
                opened.append(i)
            pg = lib.H5Gcreate2(i, b"potcar", 0, 0, 0)
            opened.append(pg)
            raw = potcar_text.encode()
            _put(lib, pg, "content", "s", raw, len(raw))
    finally:
        for g in reversed(opened):
            lib.H5Gclose(g)
        lib.H5Fclose(f)
