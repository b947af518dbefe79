"""ctypes front end of UG's data files (include/ugmg_dio.h): the
"<name>.ug.data.<xdr|asc|bin>" files of SaveData / LoadData
(np/udm/data_io.cc:408-1019).

    h, data = read(path)                    # header dict, ndata doubles
    write(path, data, mode=XDR, vd=[("sol", "u")], mgfile="...")
    per_level(data, ncomp, level_sizes)     # node-major data -> level arrays
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import ugmg

XDR, ASCII, BIN = 0, 1, 2
SCALAR, VECTOR, MULTIPLE_SCALAR = 0, 1, 2
VDMAX, NAMELEN, IDENTLEN = 100, 128, 4096


class Header(C.Structure):
    _fields_ = [("mode", C.c_int), ("version", C.c_char * NAMELEN), ("mgfile", C.c_char * NAMELEN),
                ("nparfiles", C.c_int), ("me", C.c_int), ("time", C.c_double), ("dt", C.c_double),
                ("ndt", C.c_double), ("magic_cookie", C.c_int), ("ident", C.c_char * IDENTLEN),
                ("nvd", C.c_int), ("vd_name", (C.c_char * NAMELEN) * VDMAX), ("vd_ncomp", C.c_int * VDMAX),
                ("vd_type", C.c_int * VDMAX), ("vd_comp_names", (C.c_char * NAMELEN) * VDMAX),
                ("ndata", C.c_int)]


_done = False


def _lib():
    global _done
    L = ugmg.load()
    if not _done:
        L.ugmg_dio_read.argtypes = [C.c_char_p, C.POINTER(Header), C.c_void_p, C.c_longlong]
        L.ugmg_dio_read.restype = C.c_int
        L.ugmg_dio_write.argtypes = [C.c_char_p, C.POINTER(Header), C.c_void_p]
        L.ugmg_dio_write.restype = C.c_int
        _done = True
    return L


def _as_dict(h: Header) -> dict:
    n = h.nvd
    return {"mode": h.mode, "version": h.version.decode(), "ident": h.ident.decode(),
            "mgfile": h.mgfile.decode(), "time": h.time, "dt": h.dt, "ndt": h.ndt,
            "nparfiles": h.nparfiles, "me": h.me, "magic_cookie": h.magic_cookie,
            "vd": [(h.vd_name[i].value.decode(), h.vd_ncomp[i], h.vd_type[i], h.vd_comp_names[i].value.decode())
                   for i in range(n)],
            "ndata": h.ndata}


def read(path: str):
    """(header dict, ndata doubles) of a UG data file."""
    L = _lib()
    h = Header()
    rc = L.ugmg_dio_read(path.encode(), C.byref(h), None, 0)
    if rc != 0:
        raise ValueError(f"{path}: not a UG data file ({rc})")
    data = np.zeros(h.ndata)
    rc = L.ugmg_dio_read(path.encode(), C.byref(h), data.ctypes.data, data.size)
    if rc != 0:
        raise ValueError(f"{path}: truncated node data ({rc})")
    return _as_dict(h), data


def write(path: str, data, mode: int = XDR, vd=(("sol", "u"),), mgfile: str = "saved_without_mg",
          ident: str = "---", time: float = -1.0, dt: float = -1.0, ndt: float = -1.0,
          magic_cookie: int = 0, version: str = "DATA_IO_1.7"):
    """Write node-major data as SaveData does: vd lists (name, component
    names), one character per component; type SCALAR for one component,
    MULTIPLE_SCALAR otherwise."""
    L = _lib()
    data = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)
    h = Header()
    h.mode, h.version, h.ident, h.mgfile = mode, version.encode(), ident.encode(), mgfile.encode()
    h.time, h.dt, h.ndt = time, dt, ndt
    h.nparfiles, h.me, h.magic_cookie = 1, 0, magic_cookie
    h.nvd = len(vd)
    for i, (name, comps) in enumerate(vd):
        h.vd_name[i].value = name.encode()
        h.vd_ncomp[i] = len(comps)
        h.vd_type[i] = SCALAR if len(comps) == 1 else MULTIPLE_SCALAR
        h.vd_comp_names[i].value = comps.encode()
    h.ndata = data.size
    rc = L.ugmg_dio_write(path.encode(), C.byref(h), data.ctypes.data)
    if rc != 0:
        raise ValueError(f"{path}: write failed ({rc})")


def per_level(data, ncomp: int, level_sizes, ids=None):
    """Split node-major data (ncomp doubles per node ID) into per-level
    arrays of level_sizes[l] * ncomp; ids[l] (node IDs of level l's vectors in
    list order) defaults to consecutive IDs, level after level."""
    data = np.asarray(data).reshape(-1, ncomp)
    out, start = [], 0
    for l, n in enumerate(level_sizes):
        idx = np.arange(start, start + n) if ids is None else np.asarray(ids[l])
        out.append(data[idx].reshape(-1).copy())
        start += n
    return out
