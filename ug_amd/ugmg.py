"""Python host side of libugmg (ctypes over include/ugmg.h).

Mirrors the numproc calls a UG script makes (npinit options become Params
fields; PreProcess/Iter/Solver map to preprocess/iter/solve).  Device memory
comes from torch tensors (plumbing only); every compute call goes to the HIP
library, and loading fails loudly when libugmg.so is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libugmg.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "ugmg.h")
MAXB = 4        # UGMG_MAXB: components per vector, 1..MAXB

SMOOTH_JAC, SMOOTH_GS = 0, 1
FP64, MIXED = 0, 1
BASE_EX, BASE_JAC = 0, 1
SETUP_HOST, SETUP_DEVICE = 0, 1
ERRORS = {0: "OK", -1: "E_ARG", -2: "E_HIP", -3: "E_STATE", -4: "E_NOMEM", -5: "E_SINGULAR",
          -6: "E_BUDGET", -7: "E_NODEV"}

# time_kernel kinds
K_SPMV64, K_SPMV32, K_JAC64, K_GS64, K_RESTRICT, K_INTERP, K_CYCLE, K_GS32, K_GSF64 = range(9)


class UgmgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("nu1", C.c_int), ("nu2", C.c_int), ("gamma", C.c_int), ("baselevel", C.c_int),
                ("smoother", C.c_int),
                ("damp", C.c_double * MAXB), ("lmgc_damp", C.c_double * MAXB), ("base_damp", C.c_double * MAXB),
                ("base_abslimit", C.c_double * MAXB), ("base_red", C.c_double * MAXB),
                ("base_maxiter", C.c_int), ("ex_reorder", C.c_int), ("precision", C.c_int),
                ("inner_cycles", C.c_int), ("use_graph", C.c_int), ("ex_float", C.c_int),
                ("inner_adapt", C.c_int), ("base_inv32", C.c_int),
                ("post_set", C.c_int), ("post_smoother", C.c_int), ("post_damp", C.c_double * MAXB),
                ("base_iter", C.c_int)]


class LResult(C.Structure):
    _fields_ = [("error_code", C.c_int), ("converged", C.c_int), ("first_defect", C.c_double * MAXB),
                ("last_defect", C.c_double * MAXB), ("number_of_linear_iterations", C.c_int)]


DEFECT_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double))

_lib = None


def declared_symbols(header=HEADER):
    """Function names declared in include/ugmg.h."""
    text = open(header).read()
    return sorted(set(re.findall(r"\b(ugmg_[a-z0-9_]+)\s*\(", text)))


class Elements(C.Structure):
    """ugmg_elements: element lists of one level in device memory."""
    _fields_ = [("n", C.c_int), ("nelem", C.c_int), ("nc", C.c_int), ("elem", C.c_void_p),
                ("emat", C.c_void_p), ("kref", C.c_void_p), ("scale", C.c_void_p), ("evec", C.c_void_p),
                ("skip", C.c_void_p), ("dirichlet", C.c_void_p), ("vclass", C.c_void_p),
                ("vnclass", C.c_void_p)]


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: the HIP library must be built (__graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.c_int, C.c_void_p
    sig = {
        "ugmg_params_default": ([C.POINTER(Params)], None),
        "ugmg_create": ([ip, C.POINTER(vp)], ip),
        "ugmg_destroy": ([vp], ip),
        "ugmg_last_error": ([vp], C.c_char_p),
        "ugmg_set_stream": ([vp, vp], ip),
        "ugmg_sync": ([vp], ip),
        "ugmg_set_nlevels": ([vp, ip, ip], ip),
        "ugmg_set_level": ([vp, ip, ip, vp, vp, vp, vp, vp, vp], ip),
        "ugmg_set_transfer": ([vp, ip, vp, vp, vp], ip),
        "ugmg_set_vtypes": ([vp, ip, vp], ip),
        "ugmg_set_transfer_matrix": ([vp, ip, vp, vp, vp], ip),
        "ugmg_norm_entries": ([vp], ip),
        "ugmg_set_ident": ([vp, vp, ip], ip),
        "ugmg_assemble_dirichlet": ([vp, ip, vp, vp], ip),
        "ugmg_clear_dirichlet": ([vp, ip, vp], ip),
        "ugmg_set_params": ([vp, C.POINTER(Params)], ip),
        "ugmg_preprocess": ([vp, ip], ip),
        "ugmg_postprocess": ([vp], ip),
        "ugmg_level_size": ([vp, ip, C.POINTER(ip), C.POINTER(ip)], ip),
        "ugmg_block": ([vp], ip),
        "ugmg_num_colors": ([vp, ip], ip),
        "ugmg_get_colors": ([vp, ip, vp], ip),
        "ugmg_get_pattern": ([vp, ip, vp, vp, vp], ip),
        "ugmg_dmatmul_minus": ([vp, ip, dp, dp], ip),
        "ugmg_dmatmul": ([vp, ip, dp, dp], ip),
        "ugmg_smatmul_minus": ([vp, ip, dp, dp], ip),
        "ugmg_dnrm2x": ([vp, ip, dp, dp], ip),
        "ugmg_dnrm2x_noalloc": ([vp, ip, dp, dp], ip),
        "ugmg_ddot": ([vp, ip, dp, dp, dp], ip),
        "ugmg_jacobi_step": ([vp, ip, dp, dp], ip),
        "ugmg_gs_step": ([vp, ip, dp, dp], ip),
        "ugmg_restrict": ([vp, ip, dp, dp, dp], ip),
        "ugmg_interpolate": ([vp, ip, dp, dp, dp], ip),
        "ugmg_base_solve": ([vp, dp, dp], ip),
        "ugmg_iter": ([vp, ip, dp, dp], ip),
        "ugmg_iter_host": ([vp, ip, dp, dp], ip),
        "ugmg_solve": ([vp, ip, dp, dp, dp, dp, ip, ip, C.POINTER(LResult), DEFECT_CB, vp], ip),
        "ugmg_solve_host": ([vp, ip, dp, dp, dp, dp, ip, ip, C.POINTER(LResult), DEFECT_CB, vp], ip),
        "ugmg_fullrefinelevel": ([vp], ip),
        "ugmg_solve_surface": ([vp, ip, vp, vp, dp, dp, ip, ip, C.POINTER(LResult), DEFECT_CB, vp], ip),
        "ugmg_solve_surface_host": ([vp, ip, vp, vp, dp, dp, ip, ip, C.POINTER(LResult), DEFECT_CB, vp], ip),
        "ugmg_time_kernel": ([vp, ip, ip, ip, C.POINTER(C.c_double), C.POINTER(C.c_double)], ip),
        "ugmg_solve_stats": ([vp, C.POINTER(C.c_longlong), C.POINTER(C.c_double)], ip),
        "ugmg_base_level": ([vp], ip),
        "ugmg_set_batch": ([vp, ip], ip),
        "ugmg_galerkin": ([vp, ip], ip),
        "ugmg_iter_batch": ([vp, ip, ip, dp, dp, ip], ip),
        "ugmg_bcgs_solve": ([vp, ip, ip, dp, dp, ip, dp, dp, ip, ip, ip, ip, C.POINTER(LResult), dp, ip], ip),
        "ugmg_bcgs_solve_host": ([vp, ip, ip, dp, dp, ip, dp, dp, ip, ip, ip, ip, C.POINTER(LResult), dp, ip],
                                 ip),
        "ugmg_assemble_level": ([vp, ip, C.POINTER(Elements), dp], ip),
        "ugmg_get_level": ([vp, ip, C.POINTER(ip), C.POINTER(ip), vp, vp, vp, vp], ip),
        "ugmg_color": ([ip, vp, vp, vp], ip),
        "ugmg_color_device": ([vp, ip, vp, vp, vp, C.POINTER(ip)], ip),
        "ugmg_set_setup": ([vp, ip], ip),
        "ugmg_layout_digest": ([vp, ip, vp, ip], ip),
        "ugmg_coarse_inverse": ([ip, vp, vp, vp, ip, vp, vp, C.POINTER(ip)], ip),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _vec_scalar(v):
    vals = [float(v)] * MAXB if np.isscalar(v) else [float(a) for a in v]
    return (C.c_double * MAXB)(*[vals[min(i, len(vals) - 1)] for i in range(MAXB)])


def default_params(**kw) -> Params:
    p = Params()
    load().ugmg_params_default(C.byref(p))
    for k, v in kw.items():
        if k in ("damp", "lmgc_damp", "base_damp", "base_abslimit", "base_red", "post_damp"):
            arr = getattr(p, k)
            vals = [v] * MAXB if np.isscalar(v) else list(v)
            for i in range(MAXB):
                arr[i] = vals[min(i, len(vals) - 1)]
        else:
            setattr(p, k, v)
    return p


def color(level):
    """Greedy colouring of a problems.Level (host routine of libugmg)."""
    rp = np.ascontiguousarray(level.row_ptr, np.int32)
    ci = np.ascontiguousarray(level.col_idx, np.int32)
    out = np.zeros(level.n, np.int32)
    nc = load().ugmg_color(level.n, rp.ctypes.data, ci.ctypes.data, out.ctypes.data)
    if nc < 0:
        raise UgmgError(nc, "ugmg_color")
    return nc, out


def coarse_inverse(level, reorder=1):
    """(rc, inverse, order, bandwidth) of a scalar level (host routine)."""
    rp = np.ascontiguousarray(level.row_ptr, np.int32)
    ci = np.ascontiguousarray(level.col_idx, np.int32)
    va = np.ascontiguousarray(level.vals, np.float64)
    inv = np.zeros((level.n, level.n))
    order = np.zeros(level.n, np.int32)
    bw = C.c_int()
    rc = load().ugmg_coarse_inverse(level.n, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, reorder,
                                    inv.ctypes.data, order.ctypes.data, C.byref(bw))
    return rc, inv, order, bw.value


def _ptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else t.data_ptr()


class Context:
    """One ugmg_ctx (device hierarchy + settings) on one GPU."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.ugmg_create(device, C.byref(h))
        if rc != 0:
            raise UgmgError(rc, "ugmg_create failed (no GPU?)")
        self.h = h
        self.params = default_params()
        self.problem = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.ugmg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise UgmgError(rc, self.lib.ugmg_last_error(self.h).decode())
        return rc

    # ---------------------------------------------------------------- setup
    def set_stream(self, stream=None):
        """Use a torch stream (or None for the context's own stream)."""
        s = None if stream is None else C.c_void_p(stream.cuda_stream)
        self._chk(self.lib.ugmg_set_stream(self.h, s))

    def set_hierarchy(self, problem, top=None):
        levels = problem.levels if top is None else problem.levels[:top + 1]
        self._keep = []
        self._chk(self.lib.ugmg_set_nlevels(self.h, len(levels), levels[0].b))
        self.nb = levels[0].b
        for l, lv in enumerate(levels):
            arrs = [np.ascontiguousarray(lv.row_ptr, np.int32), np.ascontiguousarray(lv.col_idx, np.int32),
                    np.ascontiguousarray(lv.vals, np.float64), np.ascontiguousarray(lv.skip, np.uint32),
                    np.ascontiguousarray(lv.vclass, np.uint8), np.ascontiguousarray(lv.vnclass, np.uint8)]
            self._chk(self.lib.ugmg_set_level(self.h, l, lv.n, *[a.ctypes.data for a in arrs]))
        for l in range(1, len(levels)):
            lv = levels[l]
            if getattr(lv, "p_blk", None) is not None:      # an interpolation matrix (transfer $M)
                arrs = [np.ascontiguousarray(lv.p_ptr, np.int32), np.ascontiguousarray(lv.p_idx, np.int32),
                        np.ascontiguousarray(lv.p_blk, np.float64)]
                self._chk(self.lib.ugmg_set_transfer_matrix(self.h, l, *[a.ctypes.data for a in arrs]))
                continue
            arrs = [np.ascontiguousarray(lv.p_ptr, np.int32), np.ascontiguousarray(lv.p_idx, np.int32),
                    np.ascontiguousarray(lv.p_w, np.float64)]
            self._chk(self.lib.ugmg_set_transfer(self.h, l, *[a.ctypes.data for a in arrs]))
        for l, lv in enumerate(levels):
            if getattr(lv, "vtype", None) is not None:
                vt = np.ascontiguousarray(lv.vtype, np.uint8)
                self._chk(self.lib.ugmg_set_vtypes(self.h, l, vt.ctypes.data))
        self.problem = problem
        self.nlevels = len(levels)

    def set_nlevels(self, nlevels, b=1):
        self._chk(self.lib.ugmg_set_nlevels(self.h, nlevels, b))
        self.nb = b
        self.nlevels = nlevels

    def set_transfer(self, level, p_ptr, p_idx, p_w):
        """ugmg_set_transfer: interpolation rows of `level` from level - 1."""
        arrs = [np.ascontiguousarray(p_ptr, np.int32), np.ascontiguousarray(p_idx, np.int32),
                np.ascontiguousarray(p_w, np.float64)]
        self._chk(self.lib.ugmg_set_transfer(self.h, level, *[a.ctypes.data for a in arrs]))

    def set_params(self, **kw):
        self.params = default_params(**kw)
        self._chk(self.lib.ugmg_set_params(self.h, C.byref(self.params)))

    def set_ident(self, ident=None):
        """VD_IDENT of the stop tests (a group leader per norm entry); None: NO_IDENT."""
        if ident is None:
            self._chk(self.lib.ugmg_set_ident(self.h, None, 0))
            return
        a = np.ascontiguousarray(ident, np.int32)
        self._chk(self.lib.ugmg_set_ident(self.h, a.ctypes.data, len(a)))

    def assemble_dirichlet(self, level, x, b):
        """AssembleDirichletBoundary on the stored level: b_j = x_j and unit rows on
        the Dirichlet components (host arrays; b is updated in place)."""
        x = np.ascontiguousarray(x, np.float64)
        assert b.dtype == np.float64 and b.flags.c_contiguous
        self._chk(self.lib.ugmg_assemble_dirichlet(self.h, level, x.ctypes.data, b.ctypes.data))

    def clear_dirichlet(self, level, x):
        """ClearDirichletValues on a list-order device vector."""
        self._chk(self.lib.ugmg_clear_dirichlet(self.h, level, _ptr(x)))

    def set_setup(self, where):
        """Where preprocess builds the level layouts: SETUP_DEVICE or SETUP_HOST."""
        self._chk(self.lib.ugmg_set_setup(self.h, where))

    def layout_digest(self, level):
        """FNV-1a digests of every array of a prepared level's device layout."""
        nd = self.lib.ugmg_layout_digest(self.h, level, None, 0)
        if nd < 0:
            self._chk(nd)
        out = np.zeros(nd, np.uint64)
        self._chk(min(0, self.lib.ugmg_layout_digest(self.h, level, out.ctypes.data, nd)))
        return out

    def color_device(self, level):
        """(colour count, colours, sweeps) of a problems.Level by the device colouring."""
        rp = np.ascontiguousarray(level.row_ptr, np.int32)
        ci = np.ascontiguousarray(level.col_idx, np.int32)
        out = np.zeros(level.n, np.int32)
        sw = C.c_int()
        nc = self.lib.ugmg_color_device(self.h, level.n, rp.ctypes.data, ci.ctypes.data, out.ctypes.data,
                                        C.byref(sw))
        if nc < 0:
            self._chk(nc)
        return nc, out, sw.value

    def preprocess(self, level=None):
        level = self.nlevels - 1 if level is None else level
        self._chk(self.lib.ugmg_preprocess(self.h, level))
        self.top = level

    def postprocess(self):
        self._chk(self.lib.ugmg_postprocess(self.h))

    def sync(self):
        self._chk(self.lib.ugmg_sync(self.h))

    def colors(self, level):
        n = self.problem.levels[level].n
        out = np.zeros(n, dtype=np.int32)
        self._chk(self.lib.ugmg_get_colors(self.h, level, out.ctypes.data))
        return out

    def pattern(self, level):
        lv = self.problem.levels[level]
        rp = np.zeros(lv.n + 1, np.int32)
        ci = np.zeros(lv.nnzb, np.int32)
        va = np.zeros(lv.nnzb * lv.b * lv.b, np.float64)
        self._chk(self.lib.ugmg_get_pattern(self.h, level, rp.ctypes.data, ci.ctypes.data, va.ctypes.data))
        return rp, ci, va

    # ------------------------------------------------------ device kernels
    def dmatmul_minus(self, level, x, y):
        self._chk(self.lib.ugmg_dmatmul_minus(self.h, level, _ptr(x), _ptr(y)))

    def dmatmul(self, level, x, y):
        self._chk(self.lib.ugmg_dmatmul(self.h, level, _ptr(x), _ptr(y)))

    def smatmul_minus(self, level, x, y):
        self._chk(self.lib.ugmg_smatmul_minus(self.h, level, _ptr(x), _ptr(y)))

    def _ncomp(self):
        """Entries of a norm: components, times the vector types."""
        r = self.lib.ugmg_norm_entries(self.h)
        if r < 0:
            self._chk(r)
        return r

    def dnrm2x(self, level, x):
        """Per-component norm (a float for scalar hierarchies)."""
        out = (C.c_double * MAXB)()
        self._chk(self.lib.ugmg_dnrm2x(self.h, level, _ptr(x), C.addressof(out)))
        b = self._ncomp()
        return out[0] if b == 1 else np.array(out[:b])

    def dnrm2x_noalloc(self, level, x):
        """dnrm2x through the entry point that allocates nothing."""
        out = (C.c_double * MAXB)()
        self._chk(self.lib.ugmg_dnrm2x_noalloc(self.h, level, _ptr(x), C.addressof(out)))
        b = self._ncomp()
        return out[0] if b == 1 else np.array(out[:b])

    def ddot(self, level, x, y):
        out = (C.c_double * MAXB)()
        self._chk(self.lib.ugmg_ddot(self.h, level, _ptr(x), _ptr(y), C.addressof(out)))
        b = self._ncomp()
        return out[0] if b == 1 else np.array(out[:b])

    def jacobi_step(self, level, x, b):
        self._chk(self.lib.ugmg_jacobi_step(self.h, level, _ptr(x), _ptr(b)))

    def gs_step(self, level, x, b):
        self._chk(self.lib.ugmg_gs_step(self.h, level, _ptr(x), _ptr(b)))

    def restrict(self, level, to_c, from_f, damp=1.0):
        """damp: one factor, or one per norm entry (VEC_SCALAR)."""
        d = _vec_scalar(damp)
        self._chk(self.lib.ugmg_restrict(self.h, level, _ptr(to_c), _ptr(from_f), C.addressof(d)))

    def interpolate(self, level, to_f, from_c, damp=1.0):
        """damp: one factor, or one per norm entry (VEC_SCALAR)."""
        d = _vec_scalar(damp)
        self._chk(self.lib.ugmg_interpolate(self.h, level, _ptr(to_f), _ptr(from_c), C.addressof(d)))

    def base_solve(self, c, b):
        self._chk(self.lib.ugmg_base_solve(self.h, _ptr(c), _ptr(b)))

    def iter(self, level, c, b):
        self._chk(self.lib.ugmg_iter(self.h, level, _ptr(c), _ptr(b)))

    def iter_host(self, level, b):
        lv = self.problem.levels[level]
        c = np.zeros(lv.n * lv.b)
        b = np.array(b, dtype=np.float64, copy=True)
        self._chk(self.lib.ugmg_iter_host(self.h, level, c.ctypes.data, b.ctypes.data))
        return c, b

    def solve(self, level, x, b, red=1e-10, abslimit=1e-300, maxiter=50, do_defect=True, record=True):
        """ls solve on device tensors x, b; returns (LResult, defects list)."""
        res = LResult()
        defects = []

        def _cb(user, it, ncomp, d):
            defects.append([d[i] for i in range(ncomp)])

        cb = DEFECT_CB(_cb) if record else DEFECT_CB()
        ab = (C.c_double * MAXB)(*[abslimit] * MAXB)
        rd = (C.c_double * MAXB)(*[red] * MAXB)
        self._chk(self.lib.ugmg_solve(self.h, level, _ptr(x), _ptr(b), C.addressof(ab), C.addressof(rd),
                                      maxiter, int(do_defect), C.byref(res), cb, None))
        return res, defects

    def solve_host(self, level, rhs, x0=None, red=1e-10, abslimit=1e-300, maxiter=50, do_defect=True):
        lv = self.problem.levels[level]
        x = np.zeros(lv.n * lv.b) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        b = np.array(rhs, dtype=np.float64, copy=True)
        res = LResult()
        defects = []

        def _cb(user, it, ncomp, d):
            defects.append([d[i] for i in range(ncomp)])

        cb = DEFECT_CB(_cb)
        ab = (C.c_double * MAXB)(*[abslimit] * MAXB)
        rd = (C.c_double * MAXB)(*[red] * MAXB)
        self._chk(self.lib.ugmg_solve_host(self.h, level, x.ctypes.data, b.ctypes.data, C.addressof(ab),
                                           C.addressof(rd), maxiter, int(do_defect), C.byref(res), cb, None))
        return x, b, res, np.array(defects)

    def solve_stats(self):
        """(V-cycles, algorithmic bytes) of the last solve."""
        cyc, by = C.c_longlong(), C.c_double()
        self._chk(self.lib.ugmg_solve_stats(self.h, C.byref(cyc), C.byref(by)))
        return cyc.value, by.value

    def base_level(self):
        r = self.lib.ugmg_base_level(self.h)
        if r < 0:
            self._chk(r)
        return r

    def fullrefinelevel(self):
        r = self.lib.ugmg_fullrefinelevel(self.h)
        if r < 0:
            self._chk(r)
        return r

    def solve_surface(self, level, xs, bs, red=1e-10, abslimit=1e-300, maxiter=50, do_defect=True,
                      record=True):
        """ugmg_solve_surface on device tensors xs[l], bs[l] for levels
        baselevel..level (None below); returns (LResult, defects list)."""
        XP = (C.c_void_p * (level + 1))(*[None if v is None else v.data_ptr() for v in xs[:level + 1]])
        BP = (C.c_void_p * (level + 1))(*[None if v is None else v.data_ptr() for v in bs[:level + 1]])
        res = LResult()
        defects = []

        def _cb(user, it, ncomp, d):
            defects.append([d[i] for i in range(ncomp)])

        cb = DEFECT_CB(_cb) if record else DEFECT_CB()
        ab = (C.c_double * MAXB)(*[abslimit] * MAXB)
        rd = (C.c_double * MAXB)(*[red] * MAXB)
        self._chk(self.lib.ugmg_solve_surface(self.h, level, XP, BP, C.addressof(ab), C.addressof(rd),
                                              maxiter, int(do_defect), C.byref(res), cb, None))
        return res, defects

    def solve_surface_host(self, level, rhs, x0=None, red=1e-10, abslimit=1e-300, maxiter=50,
                           do_defect=True):
        """ugmg_solve_surface_host: rhs[l] (and x0[l]) for levels baselevel..level
        (None below); returns (x list, b list, LResult, defects array)."""
        nb = self.nb
        bl = self.base_level()
        xs, bs = [], []
        for l in range(level + 1):
            if l < bl:
                xs.append(None)
                bs.append(None)
                continue
            n = self.problem.levels[l].n * nb
            xs.append(np.zeros(n) if x0 is None or x0[l] is None else np.array(x0[l], dtype=np.float64, copy=True))
            bs.append(np.array(rhs[l], dtype=np.float64, copy=True))
        XP = (C.c_void_p * (level + 1))(*[None if v is None else v.ctypes.data for v in xs])
        BP = (C.c_void_p * (level + 1))(*[None if v is None else v.ctypes.data for v in bs])
        res = LResult()
        defects = []

        def _cb(user, it, ncomp, d):
            defects.append([d[i] for i in range(ncomp)])

        cb = DEFECT_CB(_cb)
        ab = (C.c_double * MAXB)(*[abslimit] * MAXB)
        rd = (C.c_double * MAXB)(*[red] * MAXB)
        self._chk(self.lib.ugmg_solve_surface_host(self.h, level, XP, BP, C.addressof(ab), C.addressof(rd),
                                                   maxiter, int(do_defect), C.byref(res), cb, None))
        return xs, bs, res, np.array(defects)

    def assemble_level(self, level, n, elem, emat=None, kref=None, scale=None, evec=None, skip=None,
                       dirichlet=None, vclass=None, vnclass=None, rhs=None):
        """ugmg_assemble_level from device tensors: elem (nelem, nc) int32,
        emat (nelem, m, m) or kref (m, m) + scale (nelem,) float64, evec
        (nelem, nc, b), skip (n,) int32/uint32 bits, dirichlet (n b,),
        vclass / vnclass (n,) uint8; rhs (n b,) float64 receives the load."""
        el = Elements(n=n, nelem=elem.shape[0], nc=elem.shape[1], elem=_ptr(elem), emat=_ptr(emat),
                      kref=_ptr(kref), scale=_ptr(scale), evec=_ptr(evec), skip=_ptr(skip),
                      dirichlet=_ptr(dirichlet), vclass=_ptr(vclass), vnclass=_ptr(vnclass))
        self._chk(self.lib.ugmg_assemble_level(self.h, level, C.byref(el), _ptr(rhs)))

    def get_level(self, level):
        """(row_ptr, col_idx, vals, skip) of a level as the context holds it."""
        n, nnzb = C.c_int(), C.c_int()
        self._chk(self.lib.ugmg_get_level(self.h, level, C.byref(n), C.byref(nnzb), None, None, None, None))
        rp = np.zeros(n.value + 1, np.int32)
        ci = np.zeros(nnzb.value, np.int32)
        sk = np.zeros(n.value, np.uint32)
        b = self.nb
        va = np.zeros(nnzb.value * b * b)
        self._chk(self.lib.ugmg_get_level(self.h, level, C.byref(n), C.byref(nnzb), rp.ctypes.data,
                                          ci.ctypes.data, va.ctypes.data, sk.ctypes.data))
        return rp, ci, va, sk

    def galerkin(self, level):
        """Level-1 values := P^T A P on the GPU (then preprocess again)."""
        self._chk(self.lib.ugmg_galerkin(self.h, level))

    def get_level_values(self, level):
        """The context's values of a level (ConvertMatrix order) via the
        device pattern check (after preprocess)."""
        return self.pattern(level)[2]

    # ------------------------------------------------------- several columns
    def set_batch(self, nrhs):
        """Lay the level vectors out for nrhs columns (before preprocess)."""
        self._chk(self.lib.ugmg_set_batch(self.h, nrhs))

    def iter_batch(self, level, c, b):
        """lmgc on every column of the (nrhs, ld) device tensors c, b."""
        nrhs, ld = b.shape
        self._chk(self.lib.ugmg_iter_batch(self.h, level, nrhs, _ptr(c), _ptr(b), ld))

    def bcgs(self, level, x, b, red=1e-8, abslimit=1e-300, maxiter=100, restart=0, use_iter=True,
             do_defect=True, maxdef=None):
        """BCGSSolver on the (nrhs, ld) device tensors x, b (updated in
        place); returns (list of LResult, defects array (nrhs, maxdef))."""
        nrhs, ld = b.shape
        md = 2 * maxiter + 1 if maxdef is None else maxdef
        res = (LResult * nrhs)()
        dfs = np.zeros((nrhs, md))
        ab = (C.c_double * MAXB)(*[abslimit] * MAXB)
        rd = (C.c_double * MAXB)(*[red] * MAXB)
        self._chk(self.lib.ugmg_bcgs_solve(self.h, level, nrhs, _ptr(x), _ptr(b), ld, C.addressof(ab),
                                           C.addressof(rd), maxiter, restart, int(use_iter), int(do_defect),
                                           res, dfs.ctypes.data, md))
        return list(res), dfs

    def bcgs_host(self, level, rhs, x0=None, red=1e-8, abslimit=1e-300, maxiter=100, restart=0,
                  use_iter=True, do_defect=True):
        """BCGSSolver on host arrays rhs (nrhs, n) -> (x, b, results, defect lists)."""
        rhs = np.atleast_2d(np.asarray(rhs, dtype=np.float64))
        nrhs, n = rhs.shape
        x = np.zeros((nrhs, n)) if x0 is None else np.array(np.atleast_2d(x0), dtype=np.float64, copy=True)
        b = np.ascontiguousarray(rhs.copy())
        md = 2 * maxiter + 1
        res = (LResult * nrhs)()
        dfs = np.zeros((nrhs, md))
        ab = (C.c_double * MAXB)(*[abslimit] * MAXB)
        rd = (C.c_double * MAXB)(*[red] * MAXB)
        self._chk(self.lib.ugmg_bcgs_solve_host(self.h, level, nrhs, x.ctypes.data, b.ctypes.data, n,
                                                C.addressof(ab), C.addressof(rd), maxiter, restart,
                                                int(use_iter), int(do_defect), res, dfs.ctypes.data, md))
        seqs = [dfs[j, :min(res[j].number_of_linear_iterations + 1, md)] for j in range(nrhs)]
        return x, b, list(res), seqs

    def time_kernel(self, level, kind, reps=20):
        ms, by = C.c_double(), C.c_double()
        self._chk(self.lib.ugmg_time_kernel(self.h, level, kind, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value


# ---------------------------------------------------------------- serving

class LevelDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p), ("vals", C.c_void_p),
                ("skip", C.c_void_p), ("vclass", C.c_void_p), ("vnclass", C.c_void_p),
                ("p_ptr", C.c_void_p), ("p_idx", C.c_void_p), ("p_w", C.c_void_p)]


REQ_QUEUED, REQ_ACTIVE, REQ_DONE, REQ_FAILED = range(4)


def _load_server(L):
    if getattr(L, "_server_sigs", False):
        return
    vp, ip = C.c_void_p, C.c_int
    sig = {
        "ugmg_server_create": ([ip, C.c_double, ip, C.POINTER(vp)], ip),
        "ugmg_server_destroy": ([vp], ip),
        "ugmg_server_last_error": ([vp], C.c_char_p),
        "ugmg_server_submit": ([vp, ip, C.POINTER(LevelDesc), C.POINTER(Params), vp, C.c_double, C.c_double,
                                ip, C.POINTER(ip)], ip),
        "ugmg_server_step": ([vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip)], ip),
        "ugmg_server_status": ([vp, ip, C.POINTER(ip), C.POINTER(LResult), vp], ip),
        "ugmg_server_result": ([vp, ip, vp, ip], ip),
        "ugmg_server_stats": ([vp, C.POINTER(ip), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)], ip),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    L._server_sigs = True


class Server:
    """Continuous-batching solve server (ugmg_server_* in libugmg)."""

    def __init__(self, device=0, budget_bytes=64e9, max_slots=8):
        self.lib = load()
        _load_server(self.lib)
        h = C.c_void_p()
        rc = self.lib.ugmg_server_create(device, float(budget_bytes), max_slots, C.byref(h))
        if rc != 0:
            raise UgmgError(rc, "ugmg_server_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.ugmg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise UgmgError(rc, self.lib.ugmg_server_last_error(self.h).decode())

    def submit(self, problem, params: Params, rhs=None, red=1e-8, abslimit=1e-300, maxiter=200):
        levels = problem.levels
        descs = (LevelDesc * len(levels))()
        keep = []
        for l, lv in enumerate(levels):
            arrs = [np.ascontiguousarray(lv.row_ptr, np.int32), np.ascontiguousarray(lv.col_idx, np.int32),
                    np.ascontiguousarray(lv.vals, np.float64), np.ascontiguousarray(lv.skip, np.uint32),
                    np.ascontiguousarray(lv.vclass, np.uint8), np.ascontiguousarray(lv.vnclass, np.uint8)]
            if lv.p_ptr is not None:
                arrs += [np.ascontiguousarray(lv.p_ptr, np.int32), np.ascontiguousarray(lv.p_idx, np.int32),
                         np.ascontiguousarray(lv.p_w, np.float64)]
            keep.append(arrs)
            d = descs[l]
            d.n = lv.n
            d.row_ptr, d.col_idx, d.vals, d.skip, d.vclass, d.vnclass = [a.ctypes.data for a in arrs[:6]]
            if len(arrs) > 6:
                d.p_ptr, d.p_idx, d.p_w = [a.ctypes.data for a in arrs[6:]]
        b = np.ascontiguousarray(problem.rhs if rhs is None else rhs, np.float64)
        rid = C.c_int()
        self._chk(self.lib.ugmg_server_submit(self.h, len(levels), descs, C.byref(params), b.ctypes.data,
                                              red, abslimit, maxiter, C.byref(rid)))
        return rid.value

    def step(self):
        a, q, d = C.c_int(), C.c_int(), C.c_int()
        self._chk(self.lib.ugmg_server_step(self.h, C.byref(a), C.byref(q), C.byref(d)))
        return a.value, q.value, d.value

    def status(self, rid):
        st = C.c_int()
        res = LResult()
        times = (C.c_double * 3)()
        self._chk(self.lib.ugmg_server_status(self.h, rid, C.byref(st), C.byref(res), C.addressof(times)))
        return st.value, res, tuple(times)

    def result(self, rid, n):
        x = np.zeros(n)
        self._chk(self.lib.ugmg_server_result(self.h, rid, x.ctypes.data, n))
        return x

    def stats(self):
        """(resident groups, group iterations launched, solve iterations in them)"""
        g, la, co = C.c_int(), C.c_longlong(), C.c_longlong()
        self._chk(self.lib.ugmg_server_stats(self.h, C.byref(g), C.byref(la), C.byref(co)))
        return g.value, la.value, co.value

    def drain(self, total, sleep=0.0):
        import time
        while True:
            a, q, d = self.step()
            if d >= total:
                return
            if sleep:
                time.sleep(sleep)
