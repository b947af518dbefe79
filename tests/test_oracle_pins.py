"""CPU pins of the oracle and of the problem generator (no GPU).

The reference ships no golden vectors for this path, so the oracle is pinned
by computations that do not share its code: dense numpy/scipy products and
solves, closed-form Q1 stencils and transfer weights, the exact Galerkin
identity of nested Q1 spaces, the textbook multigrid error-propagation
operator, and SplitMix64's published output.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as O
from ug_amd import problems


# ------------------------------------------------------------------ helpers

def dense_P(P, l):
    return P.levels[l].dense_p(P.levels[l - 1].n)


def random_block_csr(n, b, density, seed, diag_boost=4.0):
    """Random block matrix in ConvertMatrix layout (diagonal block first)."""
    rng = np.random.default_rng(seed)
    rp, ci, vals = [0], [], []
    for i in range(n):
        cols = [i] + sorted(set(rng.choice(n, size=max(1, int(density * n)), replace=True)) - {i})
        for j in cols:
            blk = rng.uniform(-1, 1, (b, b))
            if j == i:
                blk += diag_boost * np.eye(b)
            ci.append(j)
            vals.append(blk.ravel())
        rp.append(len(ci))
    lv = problems.Level(n=n, b=b, row_ptr=np.array(rp, np.int32), col_idx=np.array(ci, np.int32),
                        vals=np.concatenate(vals), skip=np.zeros(n, np.uint32),
                        vclass=np.full(n, 3, np.uint8), vnclass=np.full(n, 3, np.uint8))
    return lv


# ----------------------------------------------------------------- SplitMix

def test_splitmix64_known_answer():
    # SplitMix64 (Steele, Lea, Flood 2014) from seed 0: first outputs
    # 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f
    want = [0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f]
    u = problems.splitmix_uniform(0, 3)
    assert np.array_equal(u, np.array([(z >> 11) * 2.0 ** -53 for z in want]))


# ---------------------------------------------------------------- generator

def test_q1_2d_stencil_closed_form(poisson2d):
    lv = poisson2d.levels[-1]
    N = lv.nodes_per_dir
    A = None
    for i in [N + 1, N * (N // 2) + N // 2, N * (N - 2) + N - 2]:
        cols = lv.col_idx[lv.row_ptr[i]:lv.row_ptr[i + 1]]
        vals = lv.vals[lv.row_ptr[i]:lv.row_ptr[i + 1]]
        assert cols[0] == i and vals[0] == pytest.approx(8 / 3, rel=1e-15)
        assert len(cols) == 9
        assert np.allclose(vals[1:], -1 / 3, rtol=1e-15)
    del A


def test_q1_3d_stencil_closed_form():
    P = problems.q1(3, 1, 3, problems.KAPPA_ONE, seed=1)
    lv = P.levels[-1]
    N = lv.nodes_per_dir
    h = 1.0 / (N - 1)
    i = (N // 2) * N * N + (N // 2) * N + N // 2
    cols = lv.col_idx[lv.row_ptr[i]:lv.row_ptr[i + 1]]
    vals = lv.vals[lv.row_ptr[i]:lv.row_ptr[i + 1]]
    assert len(cols) == 27 and cols[0] == i
    for c, v in zip(cols, vals):
        d = np.abs(np.array([(c % N) - (i % N), (c // N) % N - (i // N) % N, c // (N * N) - i // (N * N)]))
        k = int(d.sum())
        want = {0: 8 * h / 3, 1: 0.0, 2: -h / 6, 3: -h / 12}[k]
        assert v == pytest.approx(want, abs=1e-15)


@pytest.mark.parametrize("dim,cells,nref,kappa", [(2, 2, 4, problems.KAPPA_ONE), (3, 1, 3, problems.KAPPA_JUMP)])
def test_dirichlet_rows_and_row_sums(dim, cells, nref, kappa):
    P = problems.q1(dim, cells, nref, kappa, seed=7)
    for lv in P.levels:
        N = lv.nodes_per_dir
        for i in range(lv.n):
            c = [i % N, (i // N) % N] + ([i // (N * N)] if dim == 3 else [])
            bnd = any(x in (0, N - 1) for x in c)
            seg = slice(lv.row_ptr[i], lv.row_ptr[i + 1])
            assert lv.skip[i] == (1 if bnd else 0)
            if bnd:   # AssembleDirichletBoundary: unit row
                assert lv.vals[seg][0] == 1.0 and np.all(lv.vals[seg][1:] == 0.0)
            else:     # the stiffness matrix annihilates constants
                assert abs(lv.vals[seg].sum()) < 1e-13 * abs(lv.vals[seg][0])


@pytest.mark.parametrize("dim,cells,nref", [(2, 2, 4), (3, 1, 3)])
def test_coarse_levels_are_galerkin_products(dim, cells, nref):
    """Nested Q1 spaces: A_c = P^T A_f P on interior coarse nodes (the
    identity UG's CheckNP $G tests, np/algebra/npcheck.cc:358-396)."""
    P = problems.q1(dim, cells, nref, problems.KAPPA_JUMP, seed=11)
    for l in range(1, P.top + 1):
        Af, Ac = P.levels[l].dense(), P.levels[l - 1].dense()
        Pm = dense_P(P, l)
        G = Pm.T @ Af @ Pm
        inter = np.where(P.levels[l - 1].skip == 0)[0]
        if len(inter) == 0:
            continue
        sub = np.ix_(inter, inter)
        assert np.max(np.abs(G[sub] - Ac[sub])) < 1e-12 * np.max(np.abs(Ac[sub]))


@pytest.mark.parametrize("dim,cells,nref", [(2, 2, 3), (3, 1, 3)])
def test_interpolation_weights(dim, cells, nref):
    P = problems.q1(dim, cells, nref, problems.KAPPA_ONE, seed=2)
    for l in range(1, P.top + 1):
        lf, lc = P.levels[l], P.levels[l - 1]
        Pm = dense_P(P, l)
        assert np.allclose(Pm.sum(axis=1), 1.0, atol=0, rtol=0)          # constants
        assert set(np.unique(lf.p_w)) <= {1.0, 0.5, 0.25, 0.125}
        # (multi)linear functions are reproduced exactly
        Nf, Nc = lf.nodes_per_dir, lc.nodes_per_dir

        def coords(N, n):
            i = np.arange(n)
            c = [i % N, (i // N) % N] + ([i // (N * N)] if dim == 3 else [])
            return [x / (N - 1) for x in c]
        xf, xc = coords(Nf, lf.n), coords(Nc, lc.n)
        g = lambda x: 1.0 + 2.0 * x[0] - 3.0 * x[1] + (0.5 * x[2] if dim == 3 else 0) + 0.25 * x[0] * x[1]
        assert np.max(np.abs(Pm @ g(xc) - g(xf))) < 1e-14


def test_rhs_is_consistent_mass_times_f():
    P = problems.q1(2, 2, 3, problems.KAPPA_ONE, seed=5)
    lv = P.levels[-1]
    N = lv.nodes_per_dir
    h = 1.0 / (N - 1)
    f = problems.splitmix_uniform(5, lv.n, -1.0, 1.0)
    m1 = sl.toeplitz(np.r_[4.0, 1.0, np.zeros(N - 2)]) * h / 6.0
    m1[0, 0] = m1[-1, -1] = 2 * h / 6.0
    M = np.kron(m1, m1)
    want = M @ f
    want[lv.skip == 1] = 0.0
    assert np.max(np.abs(P.rhs - want)) < 1e-15


def test_q1_sizes_match_configs():
    P = problems.poisson2d_129()
    assert P.levels[-1].n == 16641 and P.levels[-1].nnzb == 148225
    assert [lv.n for lv in P.levels] == [9, 25, 81, 289, 1089, 4225, 16641]


# ------------------------------------------------------------ ugblas kernels

@pytest.mark.parametrize("b", [1, 3])
def test_dmatmul_against_dense(b):
    lv = random_block_csr(40, b, 0.15, seed=b)
    H = O.Hierarchy([lv])
    A = lv.dense()
    rng = np.random.default_rng(1)
    x, y = rng.uniform(-1, 1, 40 * b), rng.uniform(-1, 1, 40 * b)
    assert np.max(np.abs(O.dmatmul_minus(H, 0, x, y) - (x - A @ y))) < 1e-13
    assert np.max(np.abs(O.dmatmul(H, 0, y) - A @ y)) < 1e-13
    nrm = O.dnrm2x(H, 0, x)
    assert np.allclose(nrm, [np.linalg.norm(x[i::b]) for i in range(b)], rtol=1e-14)


# ----------------------------------------------------------------- smoothers

@pytest.mark.parametrize("b", [1, 2, 3, 4])
def test_solve_small_block_against_numpy(b):
    rng = np.random.default_rng(b)
    for _ in range(20):
        M = rng.uniform(-1, 1, (b, b)) + 3 * np.eye(b)
        r = rng.uniform(-1, 1, b)
        sol = np.zeros(b)
        rc = O.lib().orc_solve_small_block(b, M.ctypes.data, r.ctypes.data, sol.ctypes.data)
        assert rc == 0
        assert np.max(np.abs(sol - np.linalg.solve(M, r))) < 1e-13


def test_solve_small_block_pivot_and_singular():
    M = np.array([[1e-12, 1.0, 0, 0], [1.0, 1.0, 0, 0], [0, 0, 2.0, 0], [0, 0, 0, 3.0]])
    r = np.array([1.0, 2.0, 3.0, 4.0])
    sol = np.zeros(4)
    assert O.lib().orc_solve_small_block(4, M.ctypes.data, r.ctypes.data, sol.ctypes.data) == 0
    assert np.allclose(sol, np.linalg.solve(M, r), rtol=1e-12)
    S = np.array([[1.0, 2.0], [2.0, 4.0]])
    assert O.lib().orc_solve_small_block(2, S.ctypes.data, r.ctypes.data, sol.ctypes.data) == 1


def test_solve_small_block_singular_threshold_is_small_d():
    # column 0 at most 1.5e-25 after the exchange: below SMALL_D * 1e-10 =
    # DBL_EPSILON * 10 * 1e-10 = 2.2e-25 (refused), above DBL_EPSILON * 1e-10
    M = np.array([[1e-25, 1.0, 0, 0], [1.5e-25, 1.0, 1.0, 0], [0, 0, 2.0, 0], [0, 0, 0, 3.0]])
    r, sol = np.ones(4), np.zeros(4)
    assert O.lib().orc_solve_small_block(4, M.ctypes.data, r.ctypes.data, sol.ctypes.data) == 1
    M[1, 0] = 3e-25                                   # above the threshold: exchanged and solved
    assert O.lib().orc_solve_small_block(4, M.ctypes.data, r.ctypes.data, sol.ctypes.data) == 0
    np.testing.assert_allclose(M @ sol, r, rtol=1e-12)


def test_solve_small_block_size_limit():
    # MAX_SINGLE_VEC_COMP = 40: blocks of 40 components and more are refused
    rng = np.random.default_rng(39)
    M = rng.uniform(-1, 1, (39, 39)) + 40 * np.eye(39)
    r, sol = rng.uniform(-1, 1, 39), np.zeros(40)
    assert O.lib().orc_solve_small_block(39, M.ctypes.data, r.ctypes.data, sol.ctypes.data) == 0
    assert np.max(np.abs(sol[:39] - np.linalg.solve(M, r))) < 1e-13
    M40 = np.eye(40)
    assert O.lib().orc_solve_small_block(40, M40.ctypes.data, np.ones(40).ctypes.data, sol.ctypes.data) == 1


@pytest.mark.parametrize("b", [1, 3])
def test_l_jac_closed_form(b):
    lv = random_block_csr(30, b, 0.2, seed=10 + b)
    lv.vclass[[3, 7]] = 2          # copies: x = 0
    H = O.Hierarchy([lv])
    d = np.random.default_rng(2).uniform(-1, 1, 30 * b)
    rc, x = O.l_jac(H, 0, d)
    assert rc == 0
    blocks = lv.vals.reshape(-1, b, b)
    for v in range(30):
        want = np.zeros(b) if lv.vclass[v] < 3 else np.linalg.solve(blocks[lv.row_ptr[v]], d[v * b:(v + 1) * b])
        assert np.allclose(x[v * b:(v + 1) * b], want, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("b", [1, 3])
def test_l_lgs_is_lower_triangular_solve(b):
    lv = random_block_csr(30, b, 0.2, seed=20 + b)
    lv.vclass[[4, 9]] = 2
    H = O.Hierarchy([lv])
    d = np.random.default_rng(3).uniform(-1, 1, 30 * b)
    rc, x = O.l_lgs(H, 0, d)
    assert rc == 0
    act = np.repeat(lv.vclass >= 3, b)
    A = lv.dense()
    L = np.tril(A)
    # block lower triangle: keep the full diagonal blocks
    for v in range(30):
        L[v * b:(v + 1) * b, v * b:(v + 1) * b] = A[v * b:(v + 1) * b, v * b:(v + 1) * b]
    L = L[np.ix_(act, act)]
    want = np.zeros(30 * b)
    want[act] = sl.solve_triangular(L, d[act], lower=True) if b == 1 else np.linalg.solve(L, d[act])
    assert np.max(np.abs(x - want)) < 1e-12


def test_colour_ordered_lgs_is_permuted_triangular_solve(jump3d_17):
    H = O.Hierarchy(jump3d_17.levels)
    for l in (2, 3):
        lv = jump3d_17.levels[l]
        nc, col = O.greedy_color(H, l)
        rank = (col.astype(np.int64) * lv.n + np.arange(lv.n)).astype(np.int32)
        d = np.random.default_rng(l).uniform(-1, 1, lv.n)
        rc, x = O.l_lgs(H, l, d, rank=rank)
        perm = np.argsort(rank, kind="stable")
        A = lv.dense()[np.ix_(perm, perm)]
        want = np.zeros(lv.n)
        want[perm] = sl.solve_triangular(np.tril(A), d[perm], lower=True)
        assert np.max(np.abs(x - want)) < 1e-13


def test_greedy_colouring(poisson2d, jump3d_17):
    for P, ncol in ((poisson2d, 4), (jump3d_17, 8)):
        H = O.Hierarchy(P.levels)
        lv = P.levels[-1]
        nc, col = O.greedy_color(H, P.top)
        assert nc == ncol
        rows = np.repeat(np.arange(lv.n), np.diff(lv.row_ptr))
        off = lv.col_idx != rows
        assert not np.any(col[rows[off]] == col[lv.col_idx[off]])


# ------------------------------------------------------------------ transfer

def test_restrict_and_interpolate_are_masked_P_products(jump3d_17):
    P = jump3d_17
    H = O.Hierarchy(P.levels)
    rng = np.random.default_rng(4)
    for l in range(1, P.top + 1):
        lf, lc = P.levels[l], P.levels[l - 1]
        Pm = dense_P(P, l)
        bf = rng.uniform(-1, 1, lf.n)
        cc = rng.uniform(-1, 1, lc.n)
        want_r = (Pm.T @ (0.9 * bf)) * (lc.skip == 0)
        assert np.max(np.abs(O.restrict(H, l, bf, damp=0.9) - want_r)) < 1e-14
        want_i = (0.7 * (Pm @ cc)) * (lf.skip == 0)
        assert np.max(np.abs(O.interpolate(H, l, cc, damp=0.7) - want_i)) < 1e-15


# ------------------------------------------------------------------------ ex

def test_ex_band_lu_solves(jump3d_17):
    H = O.Hierarchy(jump3d_17.levels)
    for l in (1, 2, 3):
        lv = jump3d_17.levels[l]
        ex = O.Ex(H, l, reorder=1)
        assert ex.rc == 0
        order = ex.order
        assert np.array_equal(np.sort(order), np.arange(lv.n))
        A = lv.dense()
        b = np.random.default_rng(l).uniform(-1, 1, lv.n)
        rc, x, bn = ex.smooth(b)
        assert rc == 0
        assert np.max(np.abs(x - np.linalg.solve(A, b))) < 1e-12 * np.max(np.abs(x))
        assert np.max(np.abs(bn - (b - A @ x))) < 1e-12
        # bandwidth of the reordered matrix
        pos = np.empty(lv.n, int)
        pos[order] = np.arange(lv.n)
        rows = np.repeat(np.arange(lv.n), np.diff(lv.row_ptr))
        assert ex.bandwidth == np.max(np.abs(pos[rows] - pos[lv.col_idx]))


def _graph_level(rows):
    """Scalar level whose row v lists the diagonal, then rows[v] in order
    (values: 4 on the diagonal, -1 off it)."""
    n = len(rows)
    rp = np.zeros(n + 1, np.int32)
    ci, vals = [], []
    for v, nb in enumerate(rows):
        ci += [v] + list(nb)
        vals += [4.0] + [-1.0] * len(nb)
        rp[v + 1] = len(ci)
    return problems.Level(n=n, b=1, row_ptr=rp, col_idx=np.array(ci, np.int32), vals=np.array(vals),
                          skip=np.zeros(n, np.uint32), vclass=np.full(n, 3, np.uint8),
                          vnclass=np.zeros(n, np.uint8))


@pytest.mark.parametrize("rows,want,bw", [
    # 6-cycle 0-3-5-2-4-1-0: BFS from 0 ends at 2; BFS from 2 in row order
    ([[3, 1], [0, 4], [5, 4], [0, 5], [1, 2], [3, 2]], [2, 5, 4, 3, 1, 0], 2),
    # a path 4-1-0-2-3 with row 0 listing 2 before 1: the first BFS ends at 4
    ([[2, 1], [0, 4], [0, 3], [2], [1]], [4, 1, 0, 2, 3], 1),
    # the same path with row 0 listing 1 first: it ends at 3
    ([[1, 2], [0, 4], [0, 3], [2], [1]], [3, 2, 0, 1, 4], 1),
])
def test_ex_bfs_order_by_hand(rows, want, bw):
    """EXPreProcess (iter.cc:8774-8811): a breadth-first search from the first
    vector, visiting each row's off-diagonal entries in list order; a second
    one from the last vector the first reached gives the new list order.
    The expected orders are worked out by hand from that rule."""
    H = O.Hierarchy([_graph_level(rows)])
    ex = O.Ex(H, 0, reorder=1)
    assert ex.rc == 0
    assert list(ex.order) == want
    assert ex.bandwidth == bw


def test_ddot_per_component():
    import math
    rng = np.random.default_rng(4)
    for b in (1, 3):
        lv = random_block_csr(40, b, 0.2, seed=b)
        H = O.Hierarchy([lv])
        x, y = rng.uniform(-1, 1, 40 * b), rng.uniform(-1, 1, 40 * b)
        got = O.ddot(H, 0, x, y)
        want = [math.fsum(x[i::b] * y[i::b]) for i in range(b)]
        assert np.max(np.abs(got - want)) < 1e-14
        assert np.allclose(O.ddot(H, 0, x, x), O.dnrm2x(H, 0, x) ** 2, rtol=1e-14)


def test_ex_zero_pivot_fails_like_ug():
    lv = random_block_csr(6, 1, 0.3, seed=3)
    lv.vals[lv.row_ptr[0]] = 0.0
    H = O.Hierarchy([lv])
    ex = O.Ex(H, 0, reorder=0)
    assert ex.rc == 1
    rc, x, _ = ex.smooth(np.ones(6))
    assert rc == 1 and np.all(x == 0.0)


# ------------------------------------------------------------ lmgc and ls

def mg_error_operator(P, top, bl, kind, damp, nu1, nu2, rank=None, gamma=1):
    """Textbook E_l = S^nu2 (I - P (I - E_{l-1}^gamma) A_{l-1}^{-1} R A_l) S^nu1
    with E_bl = 0 (exact coarse solve), R = masked P^T, P masked by fine
    Dirichlet rows, S = I - damp M^{-1} A (M = diag or colour-ordered D+L);
    gamma coarse cycles per level (1: V-cycle, 2: W-cycle)."""
    def smoother_op(l):
        A = P.levels[l].dense()
        if kind == "jac":
            M = np.diag(np.diag(A))
        else:
            n = A.shape[0]
            H = O.Hierarchy(P.levels)
            _, col = O.greedy_color(H, l)
            rk = col.astype(np.int64) * n + np.arange(n)
            M = np.where(rk[None, :] <= rk[:, None], A, 0.0)
        return np.eye(A.shape[0]) - damp * np.linalg.solve(M, A)

    E = {bl: np.zeros((P.levels[bl].n,) * 2)}
    for l in range(bl + 1, top + 1):
        A, Ac = P.levels[l].dense(), P.levels[l - 1].dense()
        Pm = dense_P(P, l)
        Pint = Pm * (P.levels[l].skip == 0)[:, None]
        R = (Pm * (P.levels[l - 1].skip == 0)[None, :]).T
        S = smoother_op(l)
        Ec = np.linalg.matrix_power(E[l - 1], gamma)
        CG = np.eye(A.shape[0]) - Pint @ (np.eye(Ac.shape[0]) - Ec) @ np.linalg.solve(Ac, R @ A)
        E[l] = np.linalg.matrix_power(S, nu2) @ CG @ np.linalg.matrix_power(S, nu1)
    return E[top]


@pytest.mark.parametrize("kind,damp", [("jac", 0.8), ("mcgs", 1.0)])
def test_lmgc_cycle_is_textbook_vcycle(kind, damp):
    P = problems.q1(2, 2, 3, problems.KAPPA_JUMP, seed=9)     # 17x17, 4 levels
    H = O.Hierarchy(P.levels)
    top, bl = P.top, 1
    E = mg_error_operator(P, top, bl, kind, damp, 2, 2)
    A = P.levels[top].dense()
    p = O.params(smoother=O.JAC if kind == "jac" else O.MCGS, damp=damp, baselevel=bl)
    b = np.random.default_rng(0).uniform(-1, 1, P.levels[top].n)
    b[P.levels[top].skip == 1] = 0.0
    c, bn = O.lmgc_iter(H, top, p, b)
    want_c = (np.eye(A.shape[0]) - E) @ np.linalg.solve(A, b)
    assert np.max(np.abs(c - want_c)) < 1e-12 * np.max(np.abs(want_c))
    assert np.max(np.abs(bn - (b - A @ c))) < 1e-13 * np.max(np.abs(b))


@pytest.mark.parametrize("kind,damp", [("jac", 0.8), ("mcgs", 1.0)])
def test_lmgc_w_cycle_is_textbook(kind, damp):
    """gamma = 2: the coarse correction is two cycles on the coarse defect,
    E_c^2 in the recursion (5 levels, so the W shape reaches level 1 four
    times)."""
    P = problems.q1(2, 2, 4, problems.KAPPA_JUMP, seed=12)     # 33x33, 5 levels
    H = O.Hierarchy(P.levels)
    top, bl = P.top, 1
    E = mg_error_operator(P, top, bl, kind, damp, 1, 2, gamma=2)
    E1 = mg_error_operator(P, top, bl, kind, damp, 1, 2, gamma=1)
    A = P.levels[top].dense()
    p = O.params(smoother=O.JAC if kind == "jac" else O.MCGS, damp=damp, baselevel=bl, nu1=1, nu2=2, gamma=2)
    b = np.random.default_rng(1).uniform(-1, 1, P.levels[top].n)
    b[P.levels[top].skip == 1] = 0.0
    c, bn = O.lmgc_iter(H, top, p, b)
    want_c = (np.eye(A.shape[0]) - E) @ np.linalg.solve(A, b)
    v_c = (np.eye(A.shape[0]) - E1) @ np.linalg.solve(A, b)
    assert np.max(np.abs(c - want_c)) < 1e-12 * np.max(np.abs(want_c))
    assert np.max(np.abs(c - v_c)) > 1e-6 * np.max(np.abs(v_c))     # not the V-cycle
    assert np.max(np.abs(bn - (b - A @ c))) < 1e-13 * np.max(np.abs(b))


def test_ls_defect_sequence_and_stop_test(poisson2d):
    P = problems.q1(2, 2, 3, problems.KAPPA_ONE, seed=3)
    H = O.Hierarchy(P.levels)
    top = P.top
    E = mg_error_operator(P, top, 0, "jac", 0.8, 2, 2)
    A = P.levels[top].dense()
    K = A @ E @ np.linalg.inv(A)       # defect propagation b_{k+1} = A E A^{-1} b_k
    r = O.ls_solve(H, top, O.params(smoother=O.JAC, damp=0.8, baselevel=0), P.rhs, red=1e-10, maxiter=50)
    b = P.rhs.copy()
    seq = [np.linalg.norm(b)]
    for _ in range(r["nit"]):
        b = K @ b
        seq.append(np.linalg.norm(b))
    seq = np.array(seq)
    assert np.max(np.abs(r["defects"][:, 0] - seq) / seq) < 1e-8
    # stop test: first iterate below 1e-10 * first (strict <), none before
    d = r["defects"][:, 0]
    assert d[-1] < 1e-10 * d[0] and np.all(d[:-1] >= 1e-10 * d[0])
    # abslimit stops first
    r2 = O.ls_solve(H, top, O.params(smoother=O.JAC, damp=0.8, baselevel=0), P.rhs, red=1e-10,
                    abslimit=float(d[3]) * 1.0000001, maxiter=50)
    assert r2["nit"] == 3
    # maxiter exhaustion: not converged, iteration count = maxiter
    r3 = O.ls_solve(H, top, O.params(smoother=O.JAC, damp=0.8, baselevel=0), P.rhs, red=1e-10, maxiter=2)
    assert not r3["converged"] and r3["nit"] == 2


def test_base_abslimit_skips_the_coarse_solve():
    """UG's base ls stops before solving when the coarse defect is below its
    (absolute) $abslimit: then lmgc's coarse correction stays zero."""
    P = problems.q1(2, 2, 2, problems.KAPPA_ONE, seed=3)
    H = O.Hierarchy(P.levels)
    b = P.rhs.copy()
    c_solve, _ = O.lmgc_iter(H, P.top, O.params(baselevel=0, base_abslimit=1e-300), b)
    c_skip, _ = O.lmgc_iter(H, P.top, O.params(baselevel=0, base_abslimit=1e300), b)
    assert not np.allclose(c_solve, c_skip)
    E = mg_error_operator(P, P.top, 0, "jac", 0.8, 2, 2)
    # with the coarse solve skipped the cycle is smoothing + zero coarse correction
    A = P.levels[P.top].dense()
    want = (np.eye(A.shape[0]) - E) @ np.linalg.solve(A, b)
    assert np.max(np.abs(c_solve - want)) < 1e-12 * np.max(np.abs(want))


def test_poisson_convergence_rate_is_grid_independent():
    rates = []
    for nref in (4, 5, 6):
        P = problems.q1(2, 2, nref, problems.KAPPA_ONE, seed=1)
        H = O.Hierarchy(P.levels)
        r = O.ls_solve(H, P.top, O.params(smoother=O.JAC, damp=0.8, baselevel=0), P.rhs, red=1e-10)
        d = r["defects"][:, 0]
        rates.append((d[-1] / d[0]) ** (1.0 / r["nit"]))
    assert max(rates) < 0.12 and max(rates) - min(rates) < 0.03, rates


# ------------------------------------------------------------ Galerkin RAP

@pytest.mark.parametrize("which", ["q1_3d_jump", "p1_cd", "elasticity"])
def test_galerkin_is_ptap_in_the_coarse_pattern(which):
    """AssembleGalerkinByMatrix restated = dense P^T A P on the coarse
    pattern, with products outside it lumped into the diagonal."""
    if which == "q1_3d_jump":
        P = problems.q1(3, 1, 3, problems.KAPPA_JUMP, seed=5)
    elif which == "p1_cd":
        P = problems.p1_cd(3, 1e-2, 1.0, seed=5)
    else:
        P = problems.elasticity(2, seed=5)
    H = O.Hierarchy(P.levels)
    for l in range(1, P.top + 1):
        f, c = P.levels[l], P.levels[l - 1]
        b = f.b
        Pb = np.kron(f.dense_p(c.n), np.eye(b))
        G = Pb.T @ f.dense() @ Pb
        got = O.galerkin(H, l).reshape(-1, b, b)
        mask = np.zeros((c.n, c.n), dtype=bool)
        for v in range(c.n):
            mask[v, c.col_idx[c.row_ptr[v]:c.row_ptr[v + 1]]] = True
        want = np.zeros_like(got)
        for v in range(c.n):
            lump = np.zeros((b, b))
            for w in range(c.n):
                blk = G[v * b:(v + 1) * b, w * b:(w + 1) * b]
                if not mask[v, w]:
                    lump += blk
            for k in range(c.row_ptr[v], c.row_ptr[v + 1]):
                w = c.col_idx[k]
                want[k] = G[v * b:(v + 1) * b, w * b:(w + 1) * b] + (lump if w == v else 0.0)
        assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(want)), (which, l)


def test_galerkin_matches_exact_coarse_operator_inside():
    """On the Q1 jump problem the generator's coarse matrices integrate the
    fine coefficient exactly, so RAP reproduces them away from Dirichlet rows."""
    P = problems.q1(3, 1, 3, problems.KAPPA_JUMP, seed=6)
    H = O.Hierarchy(P.levels)
    for l in range(2, P.top + 1):
        c = P.levels[l - 1]
        got = O.galerkin(H, l)
        rows = np.repeat(np.arange(c.n), np.diff(c.row_ptr))
        inner = (c.skip[rows] == 0) & (c.skip[c.col_idx] == 0)
        assert np.max(np.abs(got[inner] - c.vals[inner])) < 1e-12 * np.max(np.abs(c.vals))


@pytest.mark.parametrize("ident", [[0, 0, 0], [0, 1, 1]])
def test_sc_cmp_groups_by_vd_ident(elasticity_17, ident):
    """sc_cmp with VD_IDENT (npscan.cc:1027-1057): the ls stops at the first
    iteration whose defects, summed in squares per group, fall below the
    first defect's times $red; the defect sequence itself is the ungrouped
    one."""
    P = elasticity_17
    H = O.Hierarchy(P.levels)
    plain = O.ls_solve(H, P.top, O.params(smoother=O.JAC, damp=0.6, baselevel=1), P.rhs, red=1e-6, maxiter=80)
    grp = O.ls_solve(H, P.top, O.params(smoother=O.JAC, damp=0.6, baselevel=1, ident=ident), P.rhs, red=1e-6,
                     maxiter=80)
    d = grp["defects"]
    k = grp["nit"]
    assert np.array_equal(d, plain["defects"][:k + 1]) or np.array_equal(plain["defects"], d[:plain["nit"] + 1])

    def below(x, y):
        return all(sum(x[j] ** 2 for j in range(3) if ident[j] == i) <
                   sum(y[j] ** 2 for j in range(3) if ident[j] == i) for i in range(3) if ident[i] == i)
    target = d[0] * 1e-6
    assert grp["converged"] and below(d[k], target) and not any(below(d[m], target) for m in range(k))
