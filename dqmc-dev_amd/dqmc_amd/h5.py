"""Simulation files (.h5) through libdqmc_h5.so (HDF5 1.10, ctypes).

Writes the layout util/gen_1band_unified_hub.py:929-1160 produces and
src/data.c:345-789 reads, with h5py's conventions: numpy arrays keep their
dtype, Python ints become int64 scalars, floats float64, complex128 the
compound {"r", "i"}, str a variable-length UTF-8 scalar.  A batch keeps
/params and /metadata in "<prefix>.h5.params" and links them from every
"<prefix>_<i>.h5" (h5py.ExternalLink, gen_1band_unified_hub.py:1042-1048).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libdqmc_h5.so")
I32, I64, U64, F64, C128, STR = 1, 2, 3, 4, 5, 6
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdqmc_h5.so not built (%s); run make -C dqmc-dev_amd" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.dqh5_create.argtypes = [C.c_char_p]
        L.dqh5_create.restype = C.c_int64
        L.dqh5_open.argtypes = [C.c_char_p, C.c_int]
        L.dqh5_open.restype = C.c_int64
        L.dqh5_close.argtypes = [C.c_int64]
        L.dqh5_write.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int64),
                                 C.c_void_p]
        L.dqh5_overwrite.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_void_p]
        L.dqh5_external_link.argtypes = [C.c_int64, C.c_char_p, C.c_char_p, C.c_char_p]
        L.dqh5_info.argtypes = [C.c_int64, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_int64)]
        L.dqh5_read.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_void_p, C.c_int64]
        L.dqh5_list.argtypes = [C.c_int64, C.c_char_p, C.c_char_p, C.c_int64]
        _lib = L
    return _lib


def _c(a, dtype=None):
    # C-contiguous copy that keeps 0-d arrays 0-d (np.ascontiguousarray would
    # turn them into shape (1,))
    return np.array(a, dtype=dtype, order="C", copy=True)


def _kind_of(v):
    if isinstance(v, str):
        return STR, None
    if isinstance(v, (bool, np.bool_)):
        return I64, np.asarray(int(v), dtype=np.int64)
    if isinstance(v, (int,)) and not isinstance(v, np.integer):
        return I64, np.asarray(v, dtype=np.int64)
    if isinstance(v, float):
        return F64, np.asarray(v, dtype=np.float64)
    if isinstance(v, complex):
        return C128, np.asarray(v, dtype=np.complex128)
    a = np.asarray(v)
    if a.dtype == np.int32:
        return I32, _c(a)
    if a.dtype == np.uint64:
        return U64, _c(a)
    if a.dtype.kind in "iu":
        return I64, _c(a, np.int64)
    if a.dtype == np.complex128 or a.dtype.kind == "c":
        return C128, _c(a, np.complex128)
    if a.dtype.kind == "f":
        return F64, _c(a, np.float64)
    raise TypeError("cannot store %r (dtype %s)" % (type(v), a.dtype))


class File:
    """Minimal HDF5 file handle: write(path, value), read(path), list(group)."""

    def __init__(self, path, mode="r"):
        self.L = lib()
        b = os.fsencode(path)
        if mode == "w":
            self.fid = self.L.dqh5_create(b)
        else:
            self.fid = self.L.dqh5_open(b, 1 if mode in ("a", "r+") else 0)
        if self.fid < 0:
            raise OSError("cannot open %s (mode %s)" % (path, mode))

    def close(self):
        if self.fid >= 0:
            self.L.dqh5_close(self.fid)
            self.fid = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def write(self, path, value):
        kind, arr = _kind_of(value)
        p = path.encode()
        if kind == STR:
            rc = self.L.dqh5_write(self.fid, p, STR, 0, None, C.c_char_p(value.encode()))
        else:
            dims = (C.c_int64 * max(1, arr.ndim))(*arr.shape)
            rc = self.L.dqh5_write(self.fid, p, kind, arr.ndim, dims, arr.ctypes.data_as(C.c_void_p))
        if rc:
            raise OSError("write failed: %s" % path)

    def overwrite(self, path, value):
        kind, arr = _kind_of(value)
        if self.L.dqh5_overwrite(self.fid, path.encode(), kind, arr.ctypes.data_as(C.c_void_p)):
            raise OSError("overwrite failed: %s" % path)

    def link_external(self, path, target_file, target_obj):
        if self.L.dqh5_external_link(self.fid, path.encode(), os.fsencode(target_file),
                                     target_obj.encode()):
            raise OSError("external link failed: %s" % path)

    def read(self, path):
        kind, nd = C.c_int(0), C.c_int(0)
        dims = (C.c_int64 * 8)()
        p = path.encode()
        if self.L.dqh5_info(self.fid, p, C.byref(kind), C.byref(nd), dims):
            raise KeyError(path)
        if kind.value == STR:
            buf = C.create_string_buffer(int(dims[0]) + 1)
            if self.L.dqh5_read(self.fid, p, STR, buf, len(buf)):
                raise OSError("read failed: %s" % path)
            return buf.value.decode()
        shape = tuple(int(dims[i]) for i in range(nd.value))
        dt = {I32: np.int32, I64: np.int64, U64: np.uint64, F64: np.float64, C128: np.complex128}[kind.value]
        out = np.zeros(shape, dtype=dt)
        if self.L.dqh5_read(self.fid, p, kind.value, out.ctypes.data_as(C.c_void_p), out.nbytes):
            raise OSError("read failed: %s" % path)
        return out

    def list(self, group):
        buf = C.create_string_buffer(1 << 16)
        if self.L.dqh5_list(self.fid, group.encode(), buf, len(buf)):
            raise KeyError(group)
        return [s for s in buf.value.decode().split("\n") if s]


_GROUPS = ("metadata", "params", "state", "meas_eqlt", "meas_uneqlt")


def write_sim(path, sim, params_file=None):
    """Write one simulation dict (as produced by dqmc_amd.gen) to `path`.
    With params_file, /params and /metadata go there once and `path` links to
    them (the reference's batch layout); the params file is written only if
    it does not exist yet."""
    one_file = params_file is None
    if not one_file and not os.path.exists(params_file):
        with File(params_file, "w") as f:
            for g in ("metadata", "params"):
                for k, v in sim[g].items():
                    f.write("/%s/%s" % (g, k), v)
    with File(path, "w") as f:
        if one_file:
            f.write("/params_file", os.path.basename(path))
            groups = _GROUPS
        else:
            rel = os.path.relpath(params_file, os.path.dirname(os.path.abspath(path)))
            f.write("/params_file", rel)
            f.link_external("/metadata", rel, "metadata")
            f.link_external("/params", rel, "params")
            groups = _GROUPS[2:]
        for g in groups:
            if g == "meas_uneqlt" and not sim.get(g):
                continue
            for k, v in sim[g].items():
                f.write("/%s/%s" % (g, k), v)


def read_sim(path):
    """Read a simulation file back into the dict layout of dqmc_amd.gen."""
    out = {}
    with File(path, "r") as f:
        root = f.list("/")
        for g in _GROUPS:
            if g not in root:
                continue
            out[g] = {k: f.read("/%s/%s" % (g, k)) for k in f.list("/" + g)}
    for g in ("meas_eqlt", "meas_uneqlt"):
        out.setdefault(g, {})
    return out


def create_batch_files(prefix, sims):
    """Write a batch made by dqmc_amd.gen.create_batch as "<prefix>_<i>.h5"
    sharing "<prefix>.h5.params".  Returns the list of simulation files."""
    pfile = "%s.h5.params" % prefix
    if os.path.exists(pfile):
        os.remove(pfile)
    files = []
    for i, s in enumerate(sims):
        fn = "%s_%d.h5" % (prefix, i)
        write_sim(fn, s, params_file=pfile)
        files.append(fn)
    return files
