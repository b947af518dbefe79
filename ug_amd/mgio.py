"""ctypes front end of UG's multigrid files (include/ugmg_mgio.h): the
"<name>.ug.mg.<xdr|asc|bin>" files of SaveMultiGrid / LoadMultiGrid
(gm/ugio.cc:1724, gm/gm.h:3466), rebuilt into levels.

    mg = read(path)
    mg["info"]                      # dict: dim, nlevels, version, names, counts
    lv = mg["levels"][l]            # dict: ids, xyz (n, dim), elem_tag,
                                    # elem_ptr, elem_corners, p_ptr, p_idx, p_w
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import ugmg

XDR, ASCII, BIN = 0, 1, 2


class Info(C.Structure):
    _fields_ = [("mode", C.c_int), ("version", C.c_char * 128), ("ident", C.c_char * 4096),
                ("domain", C.c_char * 128), ("multigrid", C.c_char * 128), ("format", C.c_char * 128),
                ("dim", C.c_int), ("magic_cookie", C.c_int), ("heapsize", C.c_int),
                ("nlevels", C.c_int), ("nnodes", C.c_int), ("npoints", C.c_int), ("nelements", C.c_int),
                ("nrules", C.c_int), ("ncoarse_points", C.c_int), ("ncoarse_bnd_points", C.c_int),
                ("ncoarse_elements", C.c_int)]


_done = False


def _lib():
    global _done
    L = ugmg.load()
    if not _done:
        vp, ip = C.c_void_p, C.POINTER(C.c_int)
        L.ugmg_mg_read.argtypes = [C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
        L.ugmg_mg_read.restype = C.c_int
        L.ugmg_mg_free.argtypes = [vp]
        L.ugmg_mg_free.restype = None
        L.ugmg_mg_get_info.argtypes = [vp, C.POINTER(Info)]
        L.ugmg_mg_get_info.restype = C.c_int
        L.ugmg_mg_level.argtypes = [vp, C.c_int, ip, ip, ip, ip, vp, vp, vp, vp, vp, vp, vp, vp]
        L.ugmg_mg_level.restype = C.c_int
        _done = True
    return L


class MgError(RuntimeError):
    pass


def read(path):
    """The rebuilt multigrid of a UG multigrid file (see the module doc)."""
    L = _lib()
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    rc = L.ugmg_mg_read(str(path).encode(), C.byref(h), err, 512)
    if rc != 0:
        raise MgError(f"ugmg_mg_read({path}): {err.value.decode(errors='replace')} ({rc})")
    try:
        inf = Info()
        L.ugmg_mg_get_info(h, C.byref(inf))
        info = {k: (getattr(inf, k).decode() if isinstance(getattr(inf, k), bytes) else getattr(inf, k))
                for k, _ in Info._fields_}
        dim = inf.dim
        levels = []
        for l in range(inf.nlevels):
            n, ne, nc, npp = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            if L.ugmg_mg_level(h, l, C.byref(n), C.byref(ne), C.byref(nc), C.byref(npp),
                               None, None, None, None, None, None, None, None) != 0:
                raise MgError(f"level {l}")
            ids = np.zeros(n.value, np.int32)
            xyz = np.zeros((n.value, dim))
            tag = np.zeros(ne.value, np.int32)
            eptr = np.zeros(ne.value + 1, np.int32)
            ec = np.zeros(nc.value, np.int32)
            pptr = np.zeros(n.value + 1, np.int32)
            pidx = np.zeros(npp.value, np.int32)
            pw = np.zeros(npp.value)
            L.ugmg_mg_level(h, l, None, None, None, None, ids.ctypes.data, xyz.ctypes.data, tag.ctypes.data,
                            eptr.ctypes.data, ec.ctypes.data, pptr.ctypes.data, pidx.ctypes.data,
                            pw.ctypes.data)
            levels.append({"ids": ids, "xyz": xyz, "elem_tag": tag, "elem_ptr": eptr, "elem_corners": ec,
                           "p_ptr": pptr, "p_idx": pidx, "p_w": pw})
        return {"info": info, "levels": levels}
    finally:
        L.ugmg_mg_free(h)
