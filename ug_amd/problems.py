"""Synthetic UG-layout multigrid problems (ctypes view of libugmg_problems).

The C++ generator (``ug_amd/csrc/problems.cc``) stands in for the UG
application that assembles one matrix per level; this module exposes its
arrays as numpy arrays (copies, so they outlive the C object).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field, replace

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libugmg_problems.so")

KAPPA_ONE = 0
KAPPA_JUMP = 1

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        P = C.c_void_p
        ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
        lib.ugmg_problem_q1.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(P)]
        lib.ugmg_problem_q1.restype = C.c_int
        lib.ugmg_problem_free.argtypes = [P]
        lib.ugmg_problem_nlevels.argtypes = [P]
        lib.ugmg_problem_block.argtypes = [P]
        lib.ugmg_problem_level.argtypes = [P, C.c_int, ip, ip, C.POINTER(ip), C.POINTER(ip), C.POINTER(dp),
                                           C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.POINTER(C.c_uint8)),
                                           C.POINTER(C.POINTER(C.c_uint8))]
        lib.ugmg_problem_transfer.argtypes = [P, C.c_int, ip, C.POINTER(ip), C.POINTER(ip), C.POINTER(dp)]
        lib.ugmg_problem_rhs.argtypes = [P, C.POINTER(dp)]
        lib.ugmg_problem_grid.argtypes = [P, C.c_int, ip, ip]
        lib.ugmg_splitmix_uniform.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double, dp]
        lib.ugmg_problem_p1_cd.argtypes = [C.c_int, C.c_double, C.c_double, C.c_uint64, C.POINTER(P)]
        lib.ugmg_problem_p1_cd.restype = C.c_int
        lib.ugmg_problem_elasticity.argtypes = [C.c_int, C.c_double, C.c_double, C.c_uint64, C.POINTER(P)]
        lib.ugmg_problem_elasticity.restype = C.c_int
        lib.ugmg_problem_elasticity_local.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(P)]
        lib.ugmg_problem_elasticity_local.restype = C.c_int
        lib.ugmg_problem_mesh.argtypes = [P, C.c_int, ip, C.POINTER(dp), ip, C.POINTER(ip),
                                          C.POINTER(C.POINTER(C.c_uint8))]
        lib.ugmg_problem_mesh.restype = C.c_int
        lib.ugmg_problem_theta.argtypes = [P]
        lib.ugmg_problem_theta.restype = C.c_double
        E = C.POINTER(C.POINTER(_ElementData))
        lib.ugmg_problem_q1_elements.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, E]
        lib.ugmg_problem_p1_cd_elements.argtypes = [C.c_int, C.c_double, C.c_double, C.c_uint64, E]
        lib.ugmg_problem_elasticity_elements.argtypes = [C.c_int, C.c_double, C.c_double, C.c_uint64, E]
        lib.ugmg_element_data_free.argtypes = [C.POINTER(_ElementData)]
        lib.ugmg_element_data_free.restype = None
        lib.ugmg_splitmix_uniform.restype = None
        _lib = lib
    return _lib


class _ElementData(C.Structure):
    _fields_ = [("n", C.c_int), ("b", C.c_int), ("nelem", C.c_int), ("nc", C.c_int),
                ("elem", C.POINTER(C.c_int)), ("emat", C.POINTER(C.c_double)),
                ("scale", C.POINTER(C.c_double)), ("kref", C.POINTER(C.c_double)),
                ("skip", C.POINTER(C.c_uint32))]


@dataclass
class Elements:
    """Finest-level element data of a generator (include/ugmg_problems.h):
    elem (nelem, nc) corner vectors; element matrices (nc b)^2 either emat
    (nelem, m, m) or scale (nelem,) times kref (m, m); skip (n,) VECSKIP."""
    n: int
    b: int
    elem: np.ndarray
    skip: np.ndarray
    emat: np.ndarray | None = None
    scale: np.ndarray | None = None
    kref: np.ndarray | None = None

    @property
    def nelem(self) -> int:
        return self.elem.shape[0]

    @property
    def nc(self) -> int:
        return self.elem.shape[1]


def _elements(fn, *args) -> Elements:
    lib = _load()
    h = C.POINTER(_ElementData)()
    rc = getattr(lib, fn)(*args, C.byref(h))
    if rc != 0:
        raise ValueError(f"{fn} failed ({rc})")
    try:
        d = h.contents
        m = d.nc * d.b
        el = Elements(n=d.n, b=d.b, elem=_arr(d.elem, d.nelem * d.nc, np.int32).reshape(d.nelem, d.nc),
                      skip=_arr(d.skip, d.n, np.uint32))
        if d.emat:
            el.emat = _arr(d.emat, d.nelem * m * m, np.float64).reshape(d.nelem, m, m)
        else:
            el.scale = _arr(d.scale, d.nelem, np.float64)
            el.kref = _arr(d.kref, m * m, np.float64).reshape(m, m)
    finally:
        lib.ugmg_element_data_free(h)
    return el


def q1_elements(dim: int, coarse_cells: int, nref: int, kappa: int = KAPPA_ONE, seed: int = 0) -> Elements:
    """Element data of q1(...)'s finest level: scale = kappa per element, kref = Q1 stiffness."""
    return _elements("ugmg_problem_q1_elements", dim, coarse_cells, nref, kappa, C.c_uint64(seed))


def p1_cd_elements(nref: int, eps: float = 1e-3, theta: float = float("nan"), seed: int = 0) -> Elements:
    """Element data (SUPG element matrices) of p1_cd(...)'s finest level."""
    return _elements("ugmg_problem_p1_cd_elements", nref, eps, theta, C.c_uint64(seed))


def elasticity_elements(nref: int, E: float = 1.0, nu: float = 0.3, seed: int = 0x5EED0004) -> Elements:
    """Element data (12 x 12 P1 elasticity matrices) of elasticity(...)'s finest level."""
    return _elements("ugmg_problem_elasticity_elements", nref, E, nu, C.c_uint64(seed))


def assemble_host(el: Elements, evec: np.ndarray | None = None, dirichlet: np.ndarray | None = None):
    """Element-order assembly on the host with numpy (small levels): the
    CSR (diagonal block first, then ascending columns), each entry summed
    in element order, then AssembleDirichletBoundary.  Returns
    (row_ptr, col_idx, vals (nnzb*b*b,), rhs (n*b,))."""
    n, b, nc = el.n, el.b, el.nc
    m = nc * b
    pairs = set()
    for e in range(el.nelem):
        for v in el.elem[e]:
            for w in el.elem[e]:
                pairs.add((int(v), int(w)))
    rows = [[] for _ in range(n)]
    for v, w in pairs:
        if v != w:
            rows[v].append(w)
    rp = np.zeros(n + 1, np.int32)
    cols = []
    for v in range(n):
        r = [v] + sorted(rows[v])
        cols.extend(r)
        rp[v + 1] = rp[v] + len(r)
    ci = np.array(cols, np.int32)
    pos = {}
    for v in range(n):
        for q in range(rp[v], rp[v + 1]):
            pos[(v, int(ci[q]))] = q
    vals = np.zeros((len(ci), b, b))
    rhs = np.zeros((n, b))
    for e in range(el.nelem):
        K = el.emat[e] if el.emat is not None else None
        for i in range(nc):
            v = int(el.elem[e, i])
            for j in range(nc):
                w = int(el.elem[e, j])
                blk = K[i * b:(i + 1) * b, j * b:(j + 1) * b] if K is not None else \
                    el.scale[e] * el.kref[i * b:(i + 1) * b, j * b:(j + 1) * b]
                vals[pos[(v, w)]] += blk
            if evec is not None:
                rhs[v] += evec[e, i]
    for v in range(n):
        for j in range(b):
            if (int(el.skip[v]) >> j) & 1:
                vals[rp[v]:rp[v + 1], j, :] = 0.0
                vals[rp[v], j, j] = 1.0
                rhs[v, j] = 0.0 if dirichlet is None else dirichlet[v * b + j]
    return rp, ci, vals.reshape(-1), rhs.reshape(-1)


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


@dataclass
class Level:
    n: int
    b: int
    row_ptr: np.ndarray
    col_idx: np.ndarray
    vals: np.ndarray
    skip: np.ndarray
    vclass: np.ndarray
    vnclass: np.ndarray
    nodes_per_dir: int = 0
    p_ptr: np.ndarray | None = None     # interpolation from level-1
    p_idx: np.ndarray | None = None
    p_w: np.ndarray | None = None
    vtype: np.ndarray | None = None     # VTYPE per vector (0 node, 1 edge); None: all node
    p_blk: np.ndarray | None = None     # interpolation matrix: b x b block per entry of p_idx
                                        # ([i*b+j] = coarse component i, fine component j)

    @property
    def nnzb(self) -> int:
        return int(self.row_ptr[-1])

    def dense(self) -> np.ndarray:
        """Dense matrix (n*b x n*b); small levels only."""
        b = self.b
        A = np.zeros((self.n * b, self.n * b))
        blocks = self.vals.reshape(-1, b, b)
        for v in range(self.n):
            for m in range(self.row_ptr[v], self.row_ptr[v + 1]):
                w = self.col_idx[m]
                A[v * b:(v + 1) * b, w * b:(w + 1) * b] += blocks[m]
        return A

    def dense_p(self, n_coarse: int) -> np.ndarray:
        P = np.zeros((self.n, n_coarse))
        for v in range(self.n):
            for q in range(self.p_ptr[v], self.p_ptr[v + 1]):
                P[v, self.p_idx[q]] += self.p_w[q]
        return P


@dataclass
class Mesh:
    """Tetrahedral mesh of one level: node coordinates (n, 3) in list order,
    corners (nelem, 4) in UG's corner order, kind 0 red / 1 green / 2 copy."""
    xyz: np.ndarray
    tets: np.ndarray
    kind: np.ndarray


@dataclass
class Problem:
    dim: int
    levels: list = field(default_factory=list)
    rhs: np.ndarray | None = None
    name: str = ""
    theta: float = 0.0
    meshes: list | None = None

    @property
    def top(self) -> int:
        return len(self.levels) - 1

    @property
    def b(self) -> int:
        return self.levels[0].b


def q1(dim: int, coarse_cells: int, nref: int, kappa: int = KAPPA_ONE, seed: int = 0,
       name: str = "") -> Problem:
    """Uniformly refined Q1 problem (see include/ugmg_problems.h)."""
    lib = _load()
    h = C.c_void_p()
    rc = lib.ugmg_problem_q1(dim, coarse_cells, nref, kappa, C.c_uint64(seed), C.byref(h))
    if rc != 0:
        raise ValueError(f"ugmg_problem_q1 failed ({rc})")
    return _extract(lib, h, dim, name or f"q1_{dim}d_{coarse_cells}c_{nref}r")


def p1_cd(nref: int, eps: float = 1e-3, theta: float = float("nan"), seed: int = 0, name: str = "") -> Problem:
    """2D SUPG P1 convection-diffusion (see include/ugmg_problems.h); level
    nref has (2^(nref+1)+1)^2 nodes.  theta = NaN draws the wind angle."""
    lib = _load()
    h = C.c_void_p()
    rc = lib.ugmg_problem_p1_cd(nref, eps, theta, C.c_uint64(seed), C.byref(h))
    if rc != 0:
        raise ValueError(f"ugmg_problem_p1_cd failed ({rc})")
    th = lib.ugmg_problem_theta(h)
    P = _extract(lib, h, 2, name or f"p1cd_{2 ** (nref + 1) + 1}^2")
    P.theta = th
    return P


def elasticity(nref: int, E: float = 1.0, nu: float = 0.3, seed: int = 0x5EED0004, name: str = "") -> Problem:
    """3D P1 linear elasticity on the Kuhn-split unit cube (see
    include/ugmg_problems.h): b = 3, level nref has (2^nref+1)^3 nodes."""
    lib = _load()
    h = C.c_void_p()
    rc = lib.ugmg_problem_elasticity(nref, E, nu, C.c_uint64(seed), C.byref(h))
    if rc != 0:
        raise ValueError(f"ugmg_problem_elasticity failed ({rc})")
    return _extract(lib, h, 3, name or f"elasticity_p1_{2 ** nref + 1}^3")


def _meshes(lib, h, nl):
    out = []
    for l in range(nl):
        nn, ne = C.c_int(), C.c_int()
        xyz, tets = C.POINTER(C.c_double)(), C.POINTER(C.c_int)()
        kind = C.POINTER(C.c_uint8)()
        if lib.ugmg_problem_mesh(h, l, C.byref(nn), C.byref(xyz), C.byref(ne), C.byref(tets), C.byref(kind)):
            return None
        out.append(Mesh(xyz=_arr(xyz, 3 * nn.value, np.float64).reshape(-1, 3),
                        tets=_arr(tets, 4 * ne.value, np.int32).reshape(-1, 4),
                        kind=_arr(kind, ne.value, np.uint8)))
    return out


def elasticity_local(nuni: int, depth: int, E: float = 1.0, nu: float = 0.3, name: str = "") -> Problem:
    """3D P1 linear elasticity (b = 3, clamped on x = 0) on the Kuhn-split
    unit cube refined red nuni times, then `depth` times toward the corner
    (0,0,0) with UG's green closure of tetrahedra (centre node, one son per
    side triangle) and UG's vector classes (see include/ugmg_problems.h).
    P.meshes holds each level's tetrahedra.  Right-hand sides are per level:
    surface_load()."""
    lib = _load()
    h = C.c_void_p()
    rc = lib.ugmg_problem_elasticity_local(nuni, depth, E, nu, C.byref(h))
    if rc != 0:
        raise ValueError(f"ugmg_problem_elasticity_local failed ({rc})")
    return _extract(lib, h, 3, name or f"elasticity_local_{2 ** nuni + 1}^3_{depth}l", meshes=True)


def _extract(lib, h, dim, name, meshes=False):
    try:
        nl = lib.ugmg_problem_nlevels(h)
        b = lib.ugmg_problem_block(h)
        levels = []
        for l in range(nl):
            n, nnz = C.c_int(), C.c_int()
            rp, ci = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
            v = C.POINTER(C.c_double)()
            sk = C.POINTER(C.c_uint32)()
            vc, vnc = C.POINTER(C.c_uint8)(), C.POINTER(C.c_uint8)()
            lib.ugmg_problem_level(h, l, C.byref(n), C.byref(nnz), C.byref(rp), C.byref(ci), C.byref(v),
                                   C.byref(sk), C.byref(vc), C.byref(vnc))
            d, N = C.c_int(), C.c_int()
            lib.ugmg_problem_grid(h, l, C.byref(d), C.byref(N))
            lv = Level(n=n.value, b=b,
                       row_ptr=_arr(rp, n.value + 1, np.int32), col_idx=_arr(ci, nnz.value, np.int32),
                       vals=_arr(v, nnz.value * b * b, np.float64), skip=_arr(sk, n.value, np.uint32),
                       vclass=_arr(vc, n.value, np.uint8), vnclass=_arr(vnc, n.value, np.uint8),
                       nodes_per_dir=N.value)
            if l >= 1:
                pn = C.c_int()
                pp, pi = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
                pw = C.POINTER(C.c_double)()
                lib.ugmg_problem_transfer(h, l, C.byref(pn), C.byref(pp), C.byref(pi), C.byref(pw))
                lv.p_ptr = _arr(pp, n.value + 1, np.int32)
                lv.p_idx = _arr(pi, pn.value, np.int32)
                lv.p_w = _arr(pw, pn.value, np.float64)
            levels.append(lv)
        rb = C.POINTER(C.c_double)()
        lib.ugmg_problem_rhs(h, C.byref(rb))
        rhs = _arr(rb, levels[-1].n * b, np.float64)
        ms = _meshes(lib, h, nl) if meshes else None
    finally:
        lib.ugmg_problem_free(h)
    return Problem(dim=dim, levels=levels, rhs=rhs, name=name, meshes=ms)


def splitmix_uniform(seed: int, count: int, a: float = 0.0, b: float = 1.0) -> np.ndarray:
    lib = _load()
    out = np.empty(count, dtype=np.float64)
    lib.ugmg_splitmix_uniform(C.c_uint64(seed), count, a, b, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


# Workloads named by what they are (sizes in the docstrings).
def poisson2d_129(seed: int = 0x5EED0001) -> Problem:
    """-Laplace u = f, Q1 on the unit square, 2x2 coarse quads refined 6 times: 129x129 nodes."""
    return q1(2, 2, 6, KAPPA_ONE, seed, name="poisson2d_q1_129x129")


def jump3d(nref: int = 7, seed: int = 0x5EED0002) -> Problem:
    """-div(kappa grad u) = f, Q1 hexes on the unit cube from 1 hex refined nref times
    (nref=7: 129^3 = 2,146,689 nodes), kappa = 10^U[-2,2] per finest element."""
    n = 2 ** nref + 1
    return q1(3, 1, nref, KAPPA_JUMP, seed, name=f"jump3d_q1_{n}^3")


# ------------------------------------------------------- local refinement
def _propagate(lv: Level, cls: np.ndarray) -> np.ndarray:
    """PropagateVectorClass(3) then (2) (gm/algebra.cc:4063-4080, 4145-4190):
    matrix neighbours of class-3 vectors below 3 become 2, then neighbours of
    class-2 vectors below 2 become 1 (row lists without the diagonal)."""
    cls = cls.copy()
    rp, ci = lv.row_ptr, lv.col_idx
    rows = np.repeat(np.arange(lv.n), np.diff(rp))
    off = ci != rows
    for k in (3, 2):
        src = cls[rows] == k
        dst = ci[off & src]
        hit = np.zeros(lv.n, dtype=bool)
        hit[dst] = True
        cls[hit & (cls < k)] = k - 1
    return cls


def fine_grid_dof(lv: Level) -> np.ndarray:
    """FINE_GRID_DOF (gm/algebra.cc:2908-2909): VCLASS >= 2 and VNCLASS <= 1."""
    return (lv.vclass >= 2) & (lv.vnclass <= 1)


def fullrefinelevel(P: Problem) -> int:
    """The lowest level with a FINE_GRID_DOF vector, the top level if none
    (SetSurfaceClasses, gm/algebra.cc:2885-2916)."""
    for l in range(P.top + 1):
        if fine_grid_dof(P.levels[l]).any():
            return l
    return P.top


def surface_rhs(P: Problem, bl: int, seed: int) -> list:
    """Right-hand sides f_l ~ U[-1, 1] on every vector of levels bl..top (0 on
    Dirichlet components), None below bl."""
    rng = np.random.default_rng(seed)
    out = []
    for l, lv in enumerate(P.levels):
        if l < bl:
            out.append(None)
            continue
        f = rng.uniform(-1, 1, lv.n * lv.b)
        for j in range(lv.b):
            f[j::lv.b][(lv.skip >> j) & 1 == 1] = 0.0
        out.append(f)
    return out


def surface_mask(P: Problem, l: int) -> np.ndarray:
    """The surface vectors of level l (ON_SURFACE): FINE_GRID_DOF below the
    top level, NEW_DEFECT (VCLASS >= 2) on it."""
    lv = P.levels[l]
    return fine_grid_dof(lv) if l < P.top else lv.vclass >= 2


def surface_load(P: Problem, bl: int, seed: int) -> list:
    """Right-hand sides of a surface solve: a nodal load ~U[-1,1] per
    component (SplitMix64, drawn for every vector of levels bl..top in order)
    kept on the surface vectors from FULLREFINELEVEL up, 0 elsewhere and on
    Dirichlet components; None below bl."""
    frl = fullrefinelevel(P)
    out = []
    s = seed
    for l, lv in enumerate(P.levels):
        if l < bl:
            out.append(None)
            continue
        f = splitmix_uniform(s, lv.n * lv.b, -1.0, 1.0)
        s += 1
        keep = surface_mask(P, l) if l >= frl else np.zeros(lv.n, dtype=bool)
        for j in range(lv.b):
            f[j::lv.b][~keep | ((lv.skip >> j) & 1 == 1)] = 0.0
        out.append(f)
    return out


# coupling of the block variants (symmetric positive definite); ncomp
# components take its leading ncomp x ncomp part
_M4 = np.array([[2.0, 0.5, 0.2, 0.1], [0.5, 2.0, 0.3, 0.15], [0.2, 0.3, 2.0, 0.25], [0.1, 0.15, 0.25, 2.0]])


def _p1_stiffness(xy: np.ndarray, tri: np.ndarray) -> np.ndarray:
    """P1 stiffness of -Laplace on each triangle: (ne, 3, 3)."""
    p0, p1, p2 = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
    bx = np.stack([p1[:, 1] - p2[:, 1], p2[:, 1] - p0[:, 1], p0[:, 1] - p1[:, 1]], 1)
    cx = np.stack([p2[:, 0] - p1[:, 0], p0[:, 0] - p2[:, 0], p1[:, 0] - p0[:, 0]], 1)
    area = 0.5 * np.abs((p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1]))
    return (bx[:, :, None] * bx[:, None, :] + cx[:, :, None] * cx[:, None, :]) / (4.0 * area)[:, None, None]


def _assemble_p1(xy: np.ndarray, tri: np.ndarray, ncomp: int, coupling=None):
    """Assembly over the elements in order (element matrices added as
    AddVlistMValues does), rows diagonal first then ascending columns, UG's
    Dirichlet rows (AssembleDirichletBoundary: row zeroed, diagonal block I)
    on the boundary of the unit square, every component."""
    n = len(xy)
    K = _p1_stiffness(xy, tri)
    r = np.repeat(tri, 3, axis=1).ravel()
    c = np.tile(tri, (1, 3)).ravel()
    # unique (row, col) pairs, diagonal first in each row (code 0), then the
    # columns ascending (code col + 1)
    key = r.astype(np.int64) * (n + 1) + np.where(r == c, 0, c + 1)
    uniq, inv = np.unique(key, return_inverse=True)
    vals = np.zeros(len(uniq))
    np.add.at(vals, inv, K.ravel())
    rows = uniq // (n + 1)
    code = uniq % (n + 1)
    cols = np.where(code == 0, rows, code - 1)
    diag = code == 0
    row_ptr = np.zeros(n + 1, dtype=np.int32)
    np.add.at(row_ptr, rows + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.int32)
    bnd = (np.isclose(xy[:, 0], 0) | np.isclose(xy[:, 0], 1) | np.isclose(xy[:, 1], 0) | np.isclose(xy[:, 1], 1))
    vals[bnd[rows]] = 0.0
    vals[bnd[rows] & diag] = 1.0
    if ncomp == 1:
        blocks = vals
    else:
        C = _M4[:ncomp, :ncomp] if coupling is None else np.asarray(coupling, np.float64)
        blocks = vals[:, None, None] * C[None]
        blocks[bnd[rows]] = 0.0
        blocks[bnd[rows] & diag] = np.eye(ncomp)
        blocks = blocks.ravel()
    skip = np.where(bnd, (1 << ncomp) - 1, 0).astype(np.uint32)
    return row_ptr, cols.astype(np.int32), blocks.astype(np.float64), skip


def _classes(lv: Level, tri: np.ndarray, seed_nodes: np.ndarray) -> np.ndarray:
    """SeedVectorClasses over the elements whose corners all carry node class
    3 (MinNodeClass == 3), then PropagateVectorClasses."""
    full = seed_nodes[tri].all(axis=1)
    cls = np.zeros(lv.n, dtype=np.uint8)
    cls[tri[full].ravel()] = 3
    return _propagate(lv, cls)


def p1_local(nuni: int, depth: int, ncomp: int = 1, coupling=None) -> Problem:
    """2D P1 -Laplace (ncomp = 2 .. 4: the same stiffness times the leading
    ncomp x ncomp part of a fixed 4x4 SPD coupling per entry, or times the
    ncomp x ncomp `coupling` given) on the unit
    square, coarse grid 2x2 squares split
    into 8 triangles, refined uniformly nuni times and then `depth` more
    times only toward the corner at the origin: refinement step k marks the
    red (regular) triangles inside [0, 2^-k]^2.  Each step refines as
    red-green refinement does: marked triangles red (4 children), unmarked
    triangles with one refined edge green (bisected), every other triangle
    copied (yellow).  Vector classes as UG computes them: NCLASS 3 on the
    corners of the red and green elements of a level (refine.cc:6760-6790),
    NNCLASS 3 on the corners of the coarse elements refined red or green
    (refine.cc:2890-2935), VCLASS/VNCLASS seeded from the elements whose
    corners all have class 3 and propagated over the matrix pattern
    (SetSurfaceClasses, gm/algebra.cc:2838-2930); the top level's VNCLASS
    is 0.  Level l's list: the sons of level l-1's vectors in order, then the
    edge midpoints (ascending edge); interpolation: son 1, midpoint 1/2 + 1/2.
    Level 0 holds 9 vectors; the copies keep every level on the whole
    square (far copies have class 0, off the surface)."""
    if nuni < 0 or depth < 0:
        raise ValueError("nuni, depth >= 0")
    xs = np.linspace(0.0, 1.0, 3)
    xy = np.array([(x, y) for y in xs for x in xs])
    tri = []
    for j in range(2):
        for i in range(2):
            a, b, c, d = i + 3 * j, i + 1 + 3 * j, i + 1 + 3 * (j + 1), i + 3 * (j + 1)
            tri += [(a, b, c), (a, c, d)]
    tri = np.array(tri, dtype=np.int64)
    kind = np.zeros(len(tri), dtype=np.int8)          # 0 red (regular), 1 green, 2 copy
    meshes = [(xy, tri, kind)]
    marks = []
    for step in range(nuni + depth):
        a = 1.0 if step < nuni else 0.5 ** (step - nuni + 1)
        inside = (xy[tri] <= a + 1e-12).all(axis=(1, 2))
        marked = (kind == 0) & inside
        edges_of = np.stack([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]], 1)   # (ne, 3, 2)
        ekey = np.sort(edges_of, axis=2)
        nv = len(xy)
        while True:   # closure: two or more refined edges -> red as well
            ref = np.unique((ekey[marked][:, :, 0] * nv + ekey[marked][:, :, 1]).ravel())
            hit = np.isin(ekey[:, :, 0] * nv + ekey[:, :, 1], ref).sum(axis=1)
            more = (~marked) & (hit >= 2)
            if not more.any():
                break
            if (kind[more] != 0).any():
                raise ValueError("closure reaches an irregular element")
            marked |= more
        green = (~marked) & (hit == 1)
        if (kind[green] != 0).any():
            raise ValueError("green closure of an irregular element")
        mid = {int(k): nv + i for i, k in enumerate(ref)}
        xy_new = np.vstack([xy, 0.5 * (xy[ref // nv] + xy[ref % nv])])
        tri_new, kind_new = [], []
        for e in range(len(tri)):
            p, q, r = (int(v) for v in tri[e])
            if marked[e]:
                mpq, mqr, mrp = (mid[min(u, w) * nv + max(u, w)] for u, w in ((p, q), (q, r), (r, p)))
                tri_new += [(p, mpq, mrp), (mpq, q, mqr), (mrp, mqr, r), (mpq, mqr, mrp)]
                kind_new += [0, 0, 0, 0]
            elif green[e]:
                for (u, w, o) in ((p, q, r), (q, r, p), (r, p, q)):
                    k = min(u, w) * nv + max(u, w)
                    if k in mid:
                        m = mid[k]
                        tri_new += [(u, m, o), (m, w, o)]
                        kind_new += [1, 1]
                        break
            else:
                tri_new.append((p, q, r))
                kind_new.append(2)
        marks.append((marked | green, ref))
        xy, tri, kind = xy_new, np.array(tri_new, dtype=np.int64), np.array(kind_new, dtype=np.int8)
        meshes.append((xy, tri, kind))
    levels = []
    top = len(meshes) - 1
    for l, (xy_l, tri_l, kind_l) in enumerate(meshes):
        rp, ci, vals, skip = _assemble_p1(xy_l, tri_l, ncomp, coupling)
        n = len(xy_l)
        lv = Level(n=n, b=ncomp, row_ptr=rp, col_idx=ci, vals=vals, skip=skip,
                   vclass=np.full(n, 3, np.uint8), vnclass=np.zeros(n, np.uint8))
        if l >= 1:
            nprev = len(meshes[l - 1][0])
            ref = marks[l - 1][1]
            pp = np.concatenate([np.arange(nprev + 1), nprev + 2 * np.arange(1, len(ref) + 1)]).astype(np.int32)
            pi = np.concatenate([np.arange(nprev), np.stack([ref // nprev, ref % nprev], 1).ravel()]).astype(np.int32)
            pw = np.concatenate([np.ones(nprev), np.full(2 * len(ref), 0.5)])
            lv.p_ptr, lv.p_idx, lv.p_w = pp, pi, pw
            real = np.zeros(n, dtype=bool)
            real[tri_l[kind_l < 2].ravel()] = True            # NCLASS 3
            lv.vclass = _classes(lv, tri_l, real)
        levels.append(lv)
    for l in range(top):
        xy_l, tri_l, _ = meshes[l]
        refined, _ = marks[l]
        nn3 = np.zeros(levels[l].n, dtype=bool)
        nn3[tri_l[refined].ravel()] = True                     # NNCLASS 3
        levels[l].vnclass = _classes(levels[l], tri_l, nn3)
    rhs = np.zeros(levels[-1].n * ncomp)
    return Problem(dim=2, levels=levels, rhs=rhs, name=f"p1_local_{nuni}u{depth}l_b{ncomp}" + ("" if coupling is None else "_c"))


def _p2_stiffness(xy: np.ndarray, tri: np.ndarray) -> np.ndarray:
    """P2 stiffness of -Laplace on each triangle, local dofs corners 0, 1, 2
    then the midpoints of edges (0,1), (1,2), (2,0): (ne, 6, 6), by the
    edge-midpoint rule (exact for the degree-2 integrands)."""
    p0, p1, p2 = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
    det = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1])
    area = 0.5 * np.abs(det)
    # grad lambda_i = (y_j - y_k, x_k - x_j) / det for (i, j, k) cyclic
    g = np.stack([np.stack([p1[:, 1] - p2[:, 1], p2[:, 0] - p1[:, 0]], 1),
                  np.stack([p2[:, 1] - p0[:, 1], p0[:, 0] - p2[:, 0]], 1),
                  np.stack([p0[:, 1] - p1[:, 1], p1[:, 0] - p0[:, 0]], 1)], 1) / det[:, None, None]
    K = np.zeros((len(tri), 6, 6))
    for lam in ((0.5, 0.5, 0.0), (0.0, 0.5, 0.5), (0.5, 0.0, 0.5)):
        gr = [(4 * lam[i] - 1) * g[:, i] for i in range(3)]
        for (i, j) in ((0, 1), (1, 2), (2, 0)):
            gr.append(4 * (lam[j] * g[:, i] + lam[i] * g[:, j]))
        G = np.stack(gr, 1)                                   # (ne, 6, 2)
        K += (area / 3)[:, None, None] * np.einsum("eak,ebk->eab", G, G)
    return K


def p2_uniform(nref: int, ncomp: int = 1) -> Problem:
    """2D P2 -Laplace (ncomp > 1: times the leading ncomp x ncomp part of the
    coupling of p1_local) on the unit square with node vectors (VTYPE 0) and
    edge vectors (VTYPE 1, one per edge, at its midpoint): the descriptor of
    two vector types that UG's standard transfer serves (transgrid.cc:
    461-503).  Coarse grid: 2x2 squares split into 8 triangles, nref uniform
    red refinements.  Level l's list: its node vectors (the sons of level
    l-1's nodes in order, then the midnodes of level l-1's edges), then its
    edge vectors (edges ascending by (min, max) node).  Rows diagonal first
    then ascending columns; Dirichlet on the boundary nodes and edges
    (AssembleDirichletBoundary).  Interpolation rows as
    StandardIntCorNodeVector forms them: a corner's son 1 from its father
    node, a midnode 1/2 + 1/2 from the father edge's end nodes (GNs); an edge
    vector 1/2 + 1/2 from the father objects of its two nodes (the father
    node's vector for a corner's son, the father edge's vector for a
    midnode), transgrid.cc:317-336.  Uniform refinement: VCLASS 3,
    VNCLASS 3 below the top level."""
    xs = np.linspace(0.0, 1.0, 3)
    xy = np.array([(x, y) for y in xs for x in xs])
    tri = []
    for j in range(2):
        for i in range(2):
            a, b, c, d = i + 3 * j, i + 1 + 3 * j, i + 1 + 3 * (j + 1), i + 3 * (j + 1)
            tri += [(a, b, c), (a, c, d)]
    tri = np.array(tri, dtype=np.int64)
    C = _M4[:ncomp, :ncomp]

    def edges_of(nv, tri):
        e = np.sort(np.stack([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]], 1), axis=2)
        key = np.unique((e[:, :, 0] * nv + e[:, :, 1]).ravel())
        return key                                           # ascending (min, max)

    meshes = [(xy, tri)]
    for _ in range(nref):
        nv = len(xy)
        key = edges_of(nv, tri)
        mid = {int(k): nv + i for i, k in enumerate(key)}
        xy = np.vstack([xy, 0.5 * (xy[key // nv] + xy[key % nv])])
        tn = []
        for p, q, r in tri.tolist():
            mpq, mqr, mrp = (mid[min(u, w) * nv + max(u, w)] for u, w in ((p, q), (q, r), (r, p)))
            tn += [(p, mpq, mrp), (mpq, q, mqr), (mrp, mqr, r), (mpq, mqr, mrp)]
        tri = np.array(tn, dtype=np.int64)
        meshes.append((xy, tri))
    levels = []
    prev = None
    for l, (xy_l, tri_l) in enumerate(meshes):
        nv = len(xy_l)
        key = edges_of(nv, tri_l)
        ne = len(key)
        eidx = {int(k): i for i, k in enumerate(key)}
        n = nv + ne
        dof = np.zeros((len(tri_l), 6), dtype=np.int64)
        dof[:, :3] = tri_l
        for c, (i, j) in enumerate(((0, 1), (1, 2), (2, 0))):
            k = np.minimum(tri_l[:, i], tri_l[:, j]) * nv + np.maximum(tri_l[:, i], tri_l[:, j])
            dof[:, 3 + c] = nv + np.array([eidx[int(x)] for x in k])
        K = _p2_stiffness(xy_l, tri_l)
        r = np.repeat(dof, 6, axis=1).ravel()
        c = np.tile(dof, (1, 6)).ravel()
        code = r.astype(np.int64) * (n + 1) + np.where(r == c, 0, c + 1)
        uniq, inv = np.unique(code, return_inverse=True)
        vals = np.zeros(len(uniq))
        np.add.at(vals, inv, K.ravel())
        rows = uniq // (n + 1)
        cc = uniq % (n + 1)
        cols = np.where(cc == 0, rows, cc - 1)
        diag = cc == 0
        rp = np.zeros(n + 1, dtype=np.int32)
        np.add.at(rp, rows + 1, 1)
        rp = np.cumsum(rp).astype(np.int32)
        pos = np.vstack([xy_l, 0.5 * (xy_l[key // nv] + xy_l[key % nv])])
        bnd = (np.isclose(pos[:, 0], 0) | np.isclose(pos[:, 0], 1) | np.isclose(pos[:, 1], 0) |
               np.isclose(pos[:, 1], 1))
        vals[bnd[rows]] = 0.0
        vals[bnd[rows] & diag] = 1.0
        if ncomp > 1:
            blk = vals[:, None, None] * C[None]
            blk[bnd[rows]] = 0.0
            blk[bnd[rows] & diag] = np.eye(ncomp)
            vals = blk.ravel()
        lv = Level(n=n, b=ncomp, row_ptr=rp, col_idx=cols.astype(np.int32), vals=vals.astype(np.float64),
                   skip=np.where(bnd, (1 << ncomp) - 1, 0).astype(np.uint32),
                   vclass=np.full(n, 3, np.uint8), vnclass=np.full(n, 3 if l < nref else 0, np.uint8),
                   vtype=np.concatenate([np.zeros(nv, np.uint8), np.ones(ne, np.uint8)]))
        if l >= 1:
            pnv, pkey, peidx = prev
            rows_i, rows_w = [], []
            for v in range(nv):                               # node vectors
                if v < pnv:
                    rows_i.append([v]); rows_w.append([1.0])
                else:
                    k = int(pkey[v - pnv])
                    rows_i.append([k // pnv, k % pnv]); rows_w.append([0.5, 0.5])

            def father(u):
                # a corner's son u < pnv: its father node's vector u; the midnode
                # pnv + i of coarse edge i: that edge's vector, also pnv + i
                return u
            for k in key.tolist():                            # edge vectors
                rows_i.append([father(k // nv), father(k % nv)]); rows_w.append([0.5, 0.5])
            lv.p_ptr = np.concatenate([[0], np.cumsum([len(x) for x in rows_i])]).astype(np.int32)
            lv.p_idx = np.concatenate(rows_i).astype(np.int32)
            lv.p_w = np.concatenate(rows_w).astype(np.float64)
        levels.append(lv)
        prev = (nv, key, eidx)
    rhs = np.zeros(levels[-1].n * ncomp)
    return Problem(dim=2, levels=levels, rhs=rhs, name=f"p2_{nref}r_b{ncomp}")


def with_imat(P: Problem, coupling) -> Problem:
    """P with every level's transfer as an interpolation matrix (UG's transfer
    $M, RestrictByMatrix / InterpolateCorrectionByMatrix): the block of entry q
    is p_w[q] * coupling (b x b, [i, j] = coarse component i, fine component
    j); coupling = I gives the standard weights as blocks."""
    cm = np.asarray(coupling, np.float64)
    out = []
    for l, lv in enumerate(P.levels):
        new = replace(lv)
        if l > 0:
            assert cm.shape == (lv.b, lv.b)
            new.p_blk = (np.asarray(lv.p_w)[:, None, None] * cm[None]).ravel().copy()
        out.append(new)
    return Problem(dim=P.dim, levels=out, rhs=None if P.rhs is None else np.array(P.rhs, copy=True),
                   name=P.name + "_imat", theta=P.theta, meshes=P.meshes)


def reordered(P: Problem, seed: int, newest_first: bool = True) -> Problem:
    """P with each level's vector list in a seeded random order and, with
    newest_first, every row listing its off-diagonal entries in the reverse
    of the builder's ascending order (the diagonal stays first).

    UG lists a level's vectors in creation order (gm/dlmgr.t:401-420) and
    inserts each new connection at the head of both rows' lists
    (CreateConnection, gm/algebra.cc:1042-1075), so neither the builders'
    lexicographic vector lists nor their ascending rows are UG's.  Every
    result the library computes depends on these orders only through the
    arrays it is given (row sums run in list order, the greedy colouring in
    vector-list order); this map produces such inputs.  The transfer keeps
    each fine row's entries and maps both sides; the right-hand side moves
    with the top level's vectors (rhs_new[q] = rhs[perm[q]])."""
    rng = np.random.default_rng(seed)
    perms, invs = [], []
    for lv in P.levels:
        pm = rng.permutation(lv.n).astype(np.int64)
        inv = np.empty_like(pm)
        inv[pm] = np.arange(lv.n)
        perms.append(pm)
        invs.append(inv)
    out = []
    for l, lv in enumerate(P.levels):
        pm, inv, b = perms[l], invs[l], lv.b
        bb = b * b
        blocks = lv.vals.reshape(-1, bb)
        rp = np.zeros(lv.n + 1, np.int32)
        ci, va = [], []
        for q in range(lv.n):
            r = pm[q]
            lo, hi = lv.row_ptr[r], lv.row_ptr[r + 1]
            ent = list(range(lo, hi))
            diag, off = ent[0], ent[1:]
            if newest_first:
                off = off[::-1]
            for m in [diag] + off:
                ci.append(inv[lv.col_idx[m]])
                va.append(blocks[m])
            rp[q + 1] = rp[q] + (hi - lo)
        new = Level(n=lv.n, b=b, row_ptr=rp, col_idx=np.asarray(ci, np.int32),
                    vals=np.ascontiguousarray(np.concatenate(va) if va else np.zeros(0)),
                    skip=lv.skip[pm].copy(), vclass=lv.vclass[pm].copy(), vnclass=lv.vnclass[pm].copy(),
                    vtype=None if lv.vtype is None else lv.vtype[pm].copy())
        if l > 0:
            cinv = invs[l - 1]
            pp = np.zeros(lv.n + 1, np.int32)
            pi, pw, pb = [], [], []
            bb = b * b
            for q in range(lv.n):
                r = pm[q]
                lo, hi = lv.p_ptr[r], lv.p_ptr[r + 1]
                pi.extend(cinv[lv.p_idx[lo:hi]])
                pw.extend(lv.p_w[lo:hi])
                if lv.p_blk is not None:
                    pb.append(lv.p_blk[lo * bb:hi * bb])
                pp[q + 1] = pp[q] + (hi - lo)
            new.p_ptr, new.p_idx, new.p_w = pp, np.asarray(pi, np.int32), np.asarray(pw, np.float64)
            if lv.p_blk is not None:
                new.p_blk = np.concatenate(pb) if pb else np.zeros(0)
        out.append(new)
    rhs = None
    if P.rhs is not None:
        top = P.levels[-1]
        rhs = np.asarray(P.rhs).reshape(top.n, top.b)[perms[-1]].ravel().copy()
    Q = Problem(dim=P.dim, levels=out, rhs=rhs, name=P.name + f"_reordered{seed}", theta=P.theta)
    Q.perms = perms
    return Q
