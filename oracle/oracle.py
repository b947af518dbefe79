"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline.  The product (ug_amd, libugmg) never imports it.

The oracle restates UG's solve path (see ugmg_oracle.h for the file:line of
every routine) over the flattened per-level arrays of ug_amd.problems.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

JAC, GS, MCGS = 0, 1, 2
MAXB = 4        # ORC_MAXB of ugmg_oracle.h


class OrcLevel(C.Structure):
    _fields_ = [("n", C.c_int), ("b", C.c_int),
                ("rp", C.c_void_p), ("ci", C.c_void_p), ("a", C.c_void_p),
                ("skip", C.c_void_p), ("vclass", C.c_void_p), ("vnclass", C.c_void_p),
                ("pp", C.c_void_p), ("pi", C.c_void_p), ("pw", C.c_void_p),
                ("vtype", C.c_void_p), ("ntypes", C.c_int), ("pblk", C.c_void_p)]


class OrcParams(C.Structure):
    _fields_ = [("nu1", C.c_int), ("nu2", C.c_int), ("gamma", C.c_int), ("baselevel", C.c_int),
                ("smoother", C.c_int),
                ("damp", C.c_double * MAXB), ("lmgc_damp", C.c_double * MAXB), ("base_damp", C.c_double * MAXB),
                ("base_abslimit", C.c_double * MAXB), ("base_red", C.c_double * MAXB),
                ("base_maxiter", C.c_int), ("ex_reorder", C.c_int),
                ("has_ident", C.c_int), ("ident", C.c_int * MAXB),
                ("post_set", C.c_int), ("post_smoother", C.c_int), ("post_damp", C.c_double * MAXB),
                ("base_iter", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, dp = C.c_void_p, C.c_void_p
        L.orc_dmatmul_minus.argtypes = [vp, dp, dp]
        L.orc_dmatmul.argtypes = [vp, dp, dp]
        L.orc_dnrm2x.argtypes = [vp, dp, dp]
        L.orc_ddot.argtypes = [vp, dp, dp, dp]
        L.orc_solve_small_block.argtypes = [C.c_int, dp, dp, dp]
        L.orc_l_jac.argtypes = [vp, dp, dp]
        L.orc_l_lgs.argtypes = [vp, vp, dp, dp]
        L.orc_greedy_color.argtypes = [vp, vp]
        L.orc_smoother.argtypes = [vp, C.c_int, vp, dp, dp, dp]
        L.orc_restrict.argtypes = [vp, vp, dp, dp, dp]
        L.orc_interpolate.argtypes = [vp, vp, dp, dp, dp]
        L.orc_ex_create.argtypes = [vp, C.c_int, C.POINTER(C.c_void_p)]
        L.orc_ex_free.argtypes = [vp]
        L.orc_ex_free.restype = None
        L.orc_ex_bandwidth.argtypes = [vp]
        L.orc_ex_order.argtypes = [vp]
        L.orc_ex_order.restype = C.POINTER(C.c_int)
        L.orc_ex_smoother.argtypes = [vp, vp, dp, dp, dp]
        L.orc_ls_lmgc_solve.argtypes = [vp, C.c_int, vp, dp, dp, C.c_int, dp, dp, C.c_int, dp,
                                        C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_lmgc_iter.argtypes = [vp, C.c_int, vp, dp, dp]
        L.orc_fullrefinelevel.argtypes = [vp, C.c_int]
        L.orc_ls_lmgc_solve_surface.argtypes = [vp, C.c_int, vp, vp, vp, C.c_int, dp, dp, C.c_int, dp,
                                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_galerkin.argtypes = [vp, vp, dp]
        L.orc_galerkin.restype = None
        L.orc_bcgs_solve.argtypes = [vp, C.c_int, vp, C.c_int, dp, dp, C.c_int, dp, dp, C.c_int, C.c_int, dp,
                                     C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def ntypes(levels):
    """Vector types with components: 1 + the largest VTYPE of any level."""
    vt = [lv.vtype for lv in levels if getattr(lv, "vtype", None) is not None and len(lv.vtype)]
    return 1 + max(int(np.max(v)) for v in vt) if vt else 1


class Hierarchy:
    """Keeps numpy arrays alive and exposes an OrcLevel array."""

    def __init__(self, levels):
        self._keep = []
        self.ntypes = ntypes(levels)
        arr = (OrcLevel * len(levels))()
        for i, lv in enumerate(levels):
            fields = dict(rp=np.ascontiguousarray(lv.row_ptr, np.int32),
                          ci=np.ascontiguousarray(lv.col_idx, np.int32),
                          a=np.ascontiguousarray(lv.vals, np.float64),
                          skip=np.ascontiguousarray(lv.skip, np.uint32),
                          vclass=np.ascontiguousarray(lv.vclass, np.uint8),
                          vnclass=np.ascontiguousarray(lv.vnclass, np.uint8))
            if lv.p_ptr is not None:
                fields.update(pp=np.ascontiguousarray(lv.p_ptr, np.int32),
                              pi=np.ascontiguousarray(lv.p_idx, np.int32),
                              pw=np.ascontiguousarray(lv.p_w, np.float64))
            if getattr(lv, "vtype", None) is not None:
                fields.update(vtype=np.ascontiguousarray(lv.vtype, np.uint8))
            if getattr(lv, "p_blk", None) is not None:
                fields.update(pblk=np.ascontiguousarray(lv.p_blk, np.float64))
            self._keep.append(fields)
            arr[i].n, arr[i].b, arr[i].ntypes = lv.n, lv.b, self.ntypes
            for k, v in fields.items():
                setattr(arr[i], k, v.ctypes.data)
        self.arr = arr
        self.levels = levels

    def ptr(self, l=0):
        return C.addressof(self.arr) + l * C.sizeof(OrcLevel)

    def ncomp(self, l=0):
        """Entries of a norm or VEC_SCALAR: ntypes * b."""
        return self.ntypes * self.levels[l].b


def _dv(v):
    """A VEC_SCALAR of MAXB entries: one value for all, or one per entry
    (the last repeated)."""
    vals = [float(v)] * MAXB if np.isscalar(v) else [float(a) for a in v]
    return np.array([vals[min(i, len(vals) - 1)] for i in range(MAXB)], np.float64)


def params(nu1=2, nu2=2, gamma=1, baselevel=0, smoother=JAC, damp=0.8, lmgc_damp=1.0,
           base_damp=1.0, base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, ex_reorder=1, ident=None,
           post=None, post_damp=None, base_iter=0):
    """ident: VD_IDENT per norm entry (sc_cmp's grouping), None: NO_IDENT;
    damp, lmgc_damp, base_damp: one value or one per component; post,
    post_damp: lmgc's post-smoother when it is another object than the
    pre-smoother (None: the same); base_iter: the base ls's iteration, 0 ex,
    1 jac with base_damp."""
    p = OrcParams()
    if ident is not None:
        p.has_ident = 1
        for i, v in enumerate(ident):
            p.ident[i] = v
    p.nu1, p.nu2, p.gamma, p.baselevel, p.smoother = nu1, nu2, gamma, baselevel, smoother
    dd, dl, db = _dv(damp), _dv(lmgc_damp), _dv(base_damp)
    for i in range(MAXB):
        p.damp[i] = dd[i]
        p.lmgc_damp[i] = dl[i]
        p.base_damp[i] = db[i]
        p.base_abslimit[i] = base_abslimit
        p.base_red[i] = base_red
    p.base_maxiter, p.ex_reorder, p.base_iter = base_maxiter, ex_reorder, base_iter
    if post is not None or post_damp is not None:
        p.post_set = 1
        p.post_smoother = smoother if post is None else post
        pd = _dv(damp if post_damp is None else post_damp)
        for i in range(MAXB):
            p.post_damp[i] = pd[i]
    return p


def dmatmul_minus(H, l, x, y):
    x = np.array(x, dtype=np.float64, copy=True)
    lib().orc_dmatmul_minus(H.ptr(l), x.ctypes.data, np.ascontiguousarray(y, np.float64).ctypes.data)
    return x


def dmatmul(H, l, y):
    lv = H.levels[l]
    x = np.zeros(lv.n * lv.b)
    lib().orc_dmatmul(H.ptr(l), x.ctypes.data, np.ascontiguousarray(y, np.float64).ctypes.data)
    return x


def dnrm2x(H, l, x):
    out = np.zeros(MAXB)
    lib().orc_dnrm2x(H.ptr(l), np.ascontiguousarray(x, np.float64).ctypes.data, out.ctypes.data)
    return out[:H.ncomp(l)]


def ddot(H, l, x, y):
    out = np.zeros(MAXB)
    lib().orc_ddot(H.ptr(l), np.ascontiguousarray(x, np.float64).ctypes.data,
                   np.ascontiguousarray(y, np.float64).ctypes.data, out.ctypes.data)
    return out[:H.ncomp(l)]


def l_jac(H, l, d):
    lv = H.levels[l]
    v = np.zeros(lv.n * lv.b)
    rc = lib().orc_l_jac(H.ptr(l), v.ctypes.data, np.ascontiguousarray(d, np.float64).ctypes.data)
    return rc, v


def l_lgs(H, l, d, rank=None):
    lv = H.levels[l]
    v = np.zeros(lv.n * lv.b)
    rk = None if rank is None else np.ascontiguousarray(rank, np.int32)
    rc = lib().orc_l_lgs(H.ptr(l), _p(rk), v.ctypes.data, np.ascontiguousarray(d, np.float64).ctypes.data)
    return rc, v


def greedy_color(H, l):
    col = np.zeros(H.levels[l].n, dtype=np.int32)
    nc = lib().orc_greedy_color(H.ptr(l), col.ctypes.data)
    return nc, col


def smoother(H, l, kind, damp, b, rank=None):
    """UG Smoother: returns (x, b_new)."""
    lv = H.levels[l]
    x = np.zeros(lv.n * lv.b)
    b = np.array(b, dtype=np.float64, copy=True)
    dv = _dv(damp)
    rk = None if rank is None else np.ascontiguousarray(rank, np.int32)
    rc = lib().orc_smoother(H.ptr(l), kind, _p(rk), dv.ctypes.data, x.ctypes.data, b.ctypes.data)
    if rc:
        raise RuntimeError("smoother failed")
    return x, b


def restrict(H, l, bf, damp=1.0, to_c=None):
    lc = H.levels[l - 1]
    out = np.zeros(lc.n * lc.b) if to_c is None else np.array(to_c, dtype=np.float64, copy=True)
    dv = _dv(damp)
    lib().orc_restrict(H.ptr(l), H.ptr(l - 1), out.ctypes.data, np.ascontiguousarray(bf, np.float64).ctypes.data,
                       dv.ctypes.data)
    return out


def interpolate(H, l, cc, damp=1.0):
    lf = H.levels[l]
    out = np.zeros(lf.n * lf.b)
    dv = _dv(damp)
    lib().orc_interpolate(H.ptr(l), H.ptr(l - 1), out.ctypes.data, np.ascontiguousarray(cc, np.float64).ctypes.data,
                          dv.ctypes.data)
    return out


class Ex:
    def __init__(self, H, l, reorder=1):
        self.H, self.l = H, l
        h = C.c_void_p()
        self.rc = lib().orc_ex_create(H.ptr(l), reorder, C.byref(h))
        self.h = h

    @property
    def bandwidth(self):
        return lib().orc_ex_bandwidth(self.h)

    @property
    def order(self):
        n = self.H.levels[self.l].n
        return np.ctypeslib.as_array(lib().orc_ex_order(self.h), shape=(n,)).copy()

    def smooth(self, b, damp=1.0):
        lv = self.H.levels[self.l]
        x = np.zeros(lv.n * lv.b)
        b = np.array(b, dtype=np.float64, copy=True)
        dv = _dv(damp)
        rc = lib().orc_ex_smoother(self.h, self.H.ptr(self.l), dv.ctypes.data, x.ctypes.data, b.ctypes.data)
        return rc, x, b

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ex_free(self.h)
            self.h = None


def lmgc_iter(H, top, p, b):
    lv = H.levels[top]
    c = np.zeros(lv.n * lv.b)
    b = np.array(b, dtype=np.float64, copy=True)
    rc = lib().orc_lmgc_iter(H.ptr(0), top, C.byref(p), c.ctypes.data, b.ctypes.data)
    if rc:
        raise RuntimeError(f"lmgc failed ({rc})")
    return c, b


def ls_solve(H, top, p, rhs, x0=None, red=1e-10, abslimit=1e-300, maxiter=50, do_defect=True):
    """ls + lmgc solve; returns dict(x, b, defects[nit+1, ntypes*b], nit, converged)."""
    lv = H.levels[top]
    nb = H.ncomp(top)
    x = np.zeros(lv.n * lv.b) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    b = np.array(rhs, dtype=np.float64, copy=True)
    ab, rd = _dv(abslimit), _dv(red)      # one value, or one per norm entry
    defects = np.zeros((maxiter + 1) * nb)
    nit, conv = C.c_int(), C.c_int()
    rc = lib().orc_ls_lmgc_solve(H.ptr(0), top, C.byref(p), x.ctypes.data, b.ctypes.data, int(do_defect),
                                 ab.ctypes.data, rd.ctypes.data, maxiter, defects.ctypes.data,
                                 C.byref(nit), C.byref(conv))
    if rc:
        raise RuntimeError(f"oracle solve failed ({rc})")
    k = nit.value
    return dict(x=x, b=b, defects=defects[:(k + 1) * nb].reshape(k + 1, nb), nit=k, converged=bool(conv.value))


def fullrefinelevel(H, top):
    return lib().orc_fullrefinelevel(H.ptr(0), top)


def ls_solve_surface(H, top, p, bs, xs=None, red=1e-10, abslimit=1e-300, maxiter=50, do_defect=True):
    """ls + lmgc on a locally refined hierarchy: bs[l] / xs[l] for levels
    baselevel..top (the defect recurrence runs in every level's b); returns
    dict(x=[per level], b=[per level], defects[nit+1, ntypes*b], nit, converged)."""
    nb, vb = H.ncomp(top), H.levels[top].b
    bl = min(p.baselevel, top)
    xs = [None if l < bl else (np.zeros(H.levels[l].n * vb) if xs is None or xs[l] is None
                               else np.array(xs[l], dtype=np.float64, copy=True)) for l in range(top + 1)]
    bs = [None if l < bl else np.array(bs[l], dtype=np.float64, copy=True) for l in range(top + 1)]
    XP = (C.c_void_p * (top + 1))(*[None if v is None else v.ctypes.data for v in xs])
    BP = (C.c_void_p * (top + 1))(*[None if v is None else v.ctypes.data for v in bs])
    ab = np.full(MAXB, abslimit)
    rd = np.full(MAXB, red)
    defects = np.zeros((maxiter + 1) * nb)
    nit, conv = C.c_int(), C.c_int()
    rc = lib().orc_ls_lmgc_solve_surface(H.ptr(0), top, C.byref(p), XP, BP, int(do_defect),
                                         ab.ctypes.data, rd.ctypes.data, maxiter, defects.ctypes.data,
                                         C.byref(nit), C.byref(conv))
    if rc:
        raise RuntimeError(f"oracle surface solve failed ({rc})")
    k = nit.value
    return dict(x=xs, b=bs, defects=defects[:(k + 1) * nb].reshape(k + 1, nb), nit=k,
                converged=bool(conv.value))


def bcgs_solve(H, top, p, rhs, x0=None, red=1e-8, abslimit=1e-300, maxiter=200, restart=0, use_iter=True,
               do_defect=True):
    """BCGSSolver restated; returns dict(x, b, defects[nd, b], nit, converged)."""
    lv = H.levels[top]
    nb = lv.b
    x = np.zeros(lv.n * nb) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    b = np.array(rhs, dtype=np.float64, copy=True)
    ab, rd = np.full(MAXB, abslimit), np.full(MAXB, red)
    maxdef = 2 * maxiter + 1
    defects = np.zeros(maxdef * nb)
    nd, nit, conv = C.c_int(), C.c_int(), C.c_int()
    rc = lib().orc_bcgs_solve(H.ptr(0), top, C.byref(p), int(use_iter), x.ctypes.data, b.ctypes.data,
                              int(do_defect), ab.ctypes.data, rd.ctypes.data, maxiter, restart,
                              defects.ctypes.data, maxdef, C.byref(nd), C.byref(nit), C.byref(conv))
    if rc:
        raise RuntimeError(f"oracle bcgs failed ({rc})")
    k = min(nd.value, maxdef)
    return dict(x=x, b=b, defects=defects[:k * nb].reshape(k, nb), nit=nit.value, converged=bool(conv.value))


def galerkin(H, l):
    """AssembleGalerkinByMatrix restated: values of level l-1 (in its
    pattern) from level l's matrix and interpolation."""
    lc = H.levels[l - 1]
    out = np.zeros(lc.nnzb * lc.b * lc.b)
    lib().orc_galerkin(H.ptr(l), H.ptr(l - 1), out.ctypes.data)
    return out
