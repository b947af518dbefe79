"""CPU pins for the 3x3-block (elasticity) workload and the oracle's block
paths: SolveSmallBlock-based Jacobi and Gauss-Seidel (np/algebra/ugiter.cc:
271-615, np/algebra/block.cc:104-205), per-component transfer masks
(np/algebra/transgrid.cc:117-336), ex over scalar unknowns and the lmgc
cycle, against what the mathematics fixes: rigid-body modes, strain
energies of linear fields (lambda + 2 mu, mu), P1 interpolation of linear
fields, and the textbook V-cycle operator with block smoothers.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ug_amd import problems


def coords(N):
    n = N ** 3
    iz, rem = np.divmod(np.arange(n), N * N)
    iy, ix = np.divmod(rem, N)
    h = 1.0 / (N - 1)
    return np.stack([ix * h, iy * h, iz * h], 1)


@pytest.fixture(scope="module")
def el3():
    return problems.elasticity(3, E=1.0, nu=0.3, seed=7)


def test_rigid_modes_and_energies(el3):
    lv = el3.levels[-1]
    assert lv.b == 3 and lv.n == 9 ** 3 and el3.rhs.shape == (3 * lv.n,)
    A = lv.dense()
    X = coords(lv.nodes_per_dir)
    free = np.repeat(lv.skip == 0, 3)
    for a in range(3):
        u = np.zeros((lv.n, 3))
        u[:, a] = 1.0
        assert np.max(np.abs((A @ u.ravel())[free])) < 1e-13
    for a, b in ((0, 1), (1, 2), (0, 2)):
        u = np.zeros((lv.n, 3))
        u[:, a], u[:, b] = -X[:, b], X[:, a]
        assert np.max(np.abs((A @ u.ravel())[free])) < 1e-13
    lam, mu = 0.3 / (1.3 * 0.4), 1.0 / 2.6
    u = np.zeros((lv.n, 3))
    u[:, 0] = X[:, 0]                     # eps_xx = 1, zero on the clamped face
    assert abs(u.ravel() @ A @ u.ravel() - (lam + 2 * mu)) < 1e-12
    u = np.zeros((lv.n, 3))
    u[:, 1] = X[:, 0]                     # eps_xy = 1/2
    assert abs(u.ravel() @ A @ u.ravel() - mu) < 1e-12
    # clamped rows: identity diagonal block, zero elsewhere, rhs 0
    d = lv.skip == 7
    assert np.array_equal(d, X[:, 0] == 0.0)
    rows = np.flatnonzero(np.repeat(d, 3))
    assert np.array_equal(A[rows] != 0, np.eye(3 * lv.n, dtype=bool)[rows])
    assert np.all(el3.rhs[rows] == 0.0)


def test_stencil_is_kuhn(el3):
    lv = el3.levels[-1]
    N = lv.nodes_per_dir
    for v in range(lv.n):
        cols = lv.col_idx[lv.row_ptr[v]:lv.row_ptr[v + 1]]
        assert cols[0] == v and np.all(np.diff(cols[1:]) > 0)
        z, r = divmod(v, N * N)
        y, x = divmod(r, N)
        for c in cols[1:]:
            cz, cr = divmod(int(c), N * N)
            cy, cx = divmod(cr, N)
            d = np.array([cx - x, cy - y, cz - z])
            assert np.all(d == d[np.flatnonzero(d)[0]] * (d != 0))   # +-e_S
    inner = [v for v in range(lv.n) if all(0 < t < N - 1 for t in (v % N, (v // N) % N, v // (N * N)))]
    assert np.all(np.diff(lv.row_ptr)[inner] == 15)


def test_kuhn_transfer_interpolates_linear_fields(el3):
    for l in range(1, el3.top + 1):
        f, c = el3.levels[l], el3.levels[l - 1]
        P = f.dense_p(c.n)
        Xf, Xc = coords(f.nodes_per_dir), coords(c.nodes_per_dir)
        np.testing.assert_allclose(P @ Xc, Xf, atol=1e-15)
        assert set(np.unique(f.p_w)) <= {0.5, 1.0}


def block_mg_error_operator(P, top, bl, kind, damp, nu1, nu2):
    """E = S^nu2 (I - P (I - E_c) A_c^-1 R A) S^nu1 with 3x3 blocks: S uses the
    block diagonal (Jacobi) or block lower triangle in colour order (GS); P =
    kron(scalar P, I3) with fine Dirichlet components masked, R its transpose
    with coarse Dirichlet components masked."""
    H = O.Hierarchy(P.levels)

    def smoother_op(l):
        lv = P.levels[l]
        A = lv.dense()
        n = lv.n
        if kind == "jac":
            blk = np.kron(np.eye(n), np.ones((3, 3))).astype(bool)
        else:
            _, col = O.greedy_color(H, l)
            rk = col.astype(np.int64) * n + np.arange(n)
            blk = np.kron(rk[None, :] <= rk[:, None], np.ones((3, 3))).astype(bool)
        M = np.where(blk, A, 0.0)
        return np.eye(3 * n) - damp * np.linalg.solve(M, A)

    E = {bl: np.zeros((3 * P.levels[bl].n,) * 2)}
    for l in range(bl + 1, top + 1):
        f, c = P.levels[l], P.levels[l - 1]
        A, Ac = f.dense(), c.dense()
        Pb = np.kron(f.dense_p(c.n), np.eye(3))
        fmask = ((f.skip[:, None] >> np.arange(3)) & 1).ravel() == 0
        cmask = ((c.skip[:, None] >> np.arange(3)) & 1).ravel() == 0
        Pint = Pb * fmask[:, None]
        R = (Pb * cmask[None, :]).T
        S = smoother_op(l)
        CG = np.eye(A.shape[0]) - Pint @ (np.eye(Ac.shape[0]) - E[l - 1]) @ np.linalg.solve(Ac, R @ A)
        E[l] = np.linalg.matrix_power(S, nu2) @ CG @ np.linalg.matrix_power(S, nu1)
    return E[top]


@pytest.mark.parametrize("kind,damp", [("jac", 0.6), ("mcgs", 1.0)])
def test_block_lmgc_cycle_is_textbook_vcycle(kind, damp):
    P = problems.elasticity(2, seed=11)        # 5^3 nodes, 375 unknowns
    H = O.Hierarchy(P.levels)
    top, bl = P.top, 0
    E = block_mg_error_operator(P, top, bl, kind, damp, 2, 2)
    A = P.levels[top].dense()
    p = O.params(smoother=O.JAC if kind == "jac" else O.MCGS, damp=damp, baselevel=bl)
    b = P.rhs.copy()
    c, bn = O.lmgc_iter(H, top, p, b)
    want = (np.eye(A.shape[0]) - E) @ np.linalg.solve(A, b)
    assert np.max(np.abs(c - want)) < 1e-11 * np.max(np.abs(want))
    assert np.max(np.abs(bn - (b - A @ c))) < 1e-13 * np.max(np.abs(b))


def test_block_ls_stop_test_is_per_component():
    P = problems.elasticity(3, seed=12)
    H = O.Hierarchy(P.levels)
    r = O.ls_solve(H, P.top, O.params(smoother=O.MCGS, damp=1.0, baselevel=1), P.rhs, red=1e-8, maxiter=100)
    d = r["defects"]
    assert d.shape[1] == 3 and r["converged"]
    # converged only when every component is below its own target
    assert np.all(d[-1] < 1e-8 * d[0])
    assert not np.all(d[-2] < 1e-8 * d[0])
