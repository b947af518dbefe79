"""ctypes front end of the numproc layer (include/ugmg_np.h): UG's
npcreate / npinit / npexecute for the GPU solve path.

    env = NumProcEnv(ctx)                       # ctx: ugmg.Context with its hierarchy set
    env.create("mgs", "ls")                     # npcreate mgs $c ls
    env.init("mgs", "$x sol $b rhs $A MAT $m 100 $red 1e-10 $I mgc")   # npinit
    env.execute("mgs", "$i $d $r $s $p")        # npexecute
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import ugmg

NOT_INIT, NOT_ACTIVE, ACTIVE, EXECUTABLE = range(4)

_done = False


def _lib():
    global _done
    L = ugmg.load()
    if not _done:
        vp, ip, cp = C.c_void_p, C.c_int, C.c_char_p
        sig = {
            "ugmg_np_env_create": ([vp, C.POINTER(vp)], ip),
            "ugmg_np_env_destroy": ([vp], ip),
            "ugmg_np_last_error": ([vp], cp),
            "ugmg_np_set_level": ([vp, ip], ip),
            "ugmg_np_set_vector": ([vp, cp, vp, ip], ip),
            "ugmg_np_set_level_vector": ([vp, cp, ip, vp], ip),
            "ugmg_np_create": ([vp, cp, cp], ip),
            "ugmg_np_init": ([vp, cp, cp], ip),
            "ugmg_np_execute": ([vp, cp, cp], ip),
            "ugmg_np_result": ([vp, cp, ip, C.POINTER(ugmg.LResult)], ip),
            "ugmg_np_params": ([vp, cp, C.POINTER(ugmg.Params)], ip),
            "ugmg_np_output": ([vp], cp),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _done = True
    return L


class NumProcEnv:
    """Numproc objects over one context (ctx None: configure only)."""

    def __init__(self, ctx=None):
        self.lib = _lib()
        h = C.c_void_p()
        rc = self.lib.ugmg_np_env_create(ctx.h if ctx is not None else None, C.byref(h))
        if rc != 0:
            raise ugmg.UgmgError(rc, "ugmg_np_env_create failed")
        self.h = h
        self.ctx = ctx
        self._vecs = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.ugmg_np_env_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self):
        return self.lib.ugmg_np_last_error(self.h).decode()

    def _chk(self, rc):
        if rc < 0:
            raise ugmg.UgmgError(rc, self.error())
        return rc

    def set_level(self, level):
        self._chk(self.lib.ugmg_np_set_level(self.h, level))

    def vector(self, name, array):
        """Register a copy of `array` as the named vector; returns the copy,
        which npexecute updates in place (the caller's array is untouched)."""
        a = np.array(array, dtype=np.float64, copy=True, order="C")
        ncols = a.shape[0] if a.ndim == 2 else 1
        self._vecs[name] = a
        self._chk(self.lib.ugmg_np_set_vector(self.h, name.encode(), a.ctypes.data, ncols))
        return a

    def level_vector(self, name, level, array):
        """Register a copy of `array` as the named vector's coarser level
        `level` (locally refined hierarchies); returns the copy."""
        a = np.array(array, dtype=np.float64, copy=True, order="C")
        self._vecs[(name, level)] = a
        self._chk(self.lib.ugmg_np_set_level_vector(self.h, name.encode(), level, a.ctypes.data))
        return a

    def create(self, name, cls):
        return self._chk(self.lib.ugmg_np_create(self.h, name.encode(), cls.encode()))

    def init(self, name, options=""):
        """Returns the numproc state (NOT_ACTIVE, ACTIVE, EXECUTABLE)."""
        return self._chk(self.lib.ugmg_np_init(self.h, name.encode(), options.encode()))

    def execute(self, name, flags):
        return self._chk(self.lib.ugmg_np_execute(self.h, name.encode(), flags.encode()))

    def result(self, name, col=0):
        r = ugmg.LResult()
        self._chk(self.lib.ugmg_np_result(self.h, name.encode(), col, C.byref(r)))
        return r

    def output(self):
        """The display text of the solves since the last call (UG's PCR lines)."""
        return self.lib.ugmg_np_output(self.h).decode()

    def params(self, name):
        p = ugmg.Params()
        self._chk(self.lib.ugmg_np_params(self.h, name.encode(), C.byref(p)))
        return p
