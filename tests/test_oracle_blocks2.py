"""CPU pins of the oracle on 2x2 blocks (b = 2: SolveSmallBlock's n = 2
form, np/algebra/block.cc:119-130): the 2D P1 coupled system of
problems.p1_local(.., ncomp=2) is the scalar stiffness times an SPD 2x2
coupling, and one lmgc cycle with block Jacobi or block MC-GS equals the
textbook V-cycle operator built densely with kron(P, I2) transfers masked
per Dirichlet component.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ug_amd import problems


def _ops(P, l, kind, damp, H):
    lv = P.levels[l]
    b, n = lv.b, lv.n
    A = lv.dense()
    if kind == "jac":
        blk = np.kron(np.eye(n), np.ones((b, b))).astype(bool)
    else:
        _, col = O.greedy_color(H, l)
        rk = col.astype(np.int64) * n + np.arange(n)
        blk = np.kron(rk[None, :] <= rk[:, None], np.ones((b, b))).astype(bool)
    M = np.where(blk, A, 0.0)
    return A, np.eye(b * n) - damp * np.linalg.solve(M, A)


def mg_error(P, top, bl, kind, damp, nu1, nu2):
    H = O.Hierarchy(P.levels)
    b = P.levels[0].b
    E = {bl: np.zeros((b * P.levels[bl].n,) * 2)}
    for l in range(bl + 1, top + 1):
        f, c = P.levels[l], P.levels[l - 1]
        A, S = _ops(P, l, kind, damp, H)
        Ac = c.dense()
        Pb = np.kron(f.dense_p(c.n), np.eye(b))
        fmask = ((f.skip[:, None] >> np.arange(b)) & 1).ravel() == 0
        cmask = ((c.skip[:, None] >> np.arange(b)) & 1).ravel() == 0
        CG = np.eye(A.shape[0]) - (Pb * fmask[:, None]) @ (np.eye(Ac.shape[0]) - E[l - 1]) @ \
            np.linalg.solve(Ac, (Pb * cmask[None, :]).T @ A)
        E[l] = np.linalg.matrix_power(S, nu2) @ CG @ np.linalg.matrix_power(S, nu1)
    return E[top]


def test_coupled_system_is_stiffness_times_coupling():
    P = problems.p1_local(2, 0, 2)
    s = problems.p1_local(2, 0, 1)
    for a, b in zip(P.levels, s.levels):
        assert a.b == 2
        K = np.kron(b.dense(), np.array([[2.0, 0.5], [0.5, 2.0]]))
        free = np.repeat(b.skip == 0, 2)
        assert np.max(np.abs(a.dense()[free] - K[free])) < 1e-14
        assert np.array_equal(a.skip, np.where(b.skip != 0, 3, 0))


@pytest.mark.parametrize("kind,damp", [("jac", 0.6), ("mcgs", 1.0)])
def test_block2_lmgc_cycle_is_textbook_vcycle(kind, damp):
    P = problems.p1_local(3, 0, 2)
    H = O.Hierarchy(P.levels)
    top, bl = P.top, 1
    E = mg_error(P, top, bl, kind, damp, 2, 2)
    A = P.levels[top].dense()
    n2 = A.shape[0]
    rhs = np.random.default_rng(4).uniform(-1, 1, n2)
    rhs[np.repeat(P.levels[top].skip != 0, 2)] = 0.0
    p = O.params(smoother=O.JAC if kind == "jac" else O.MCGS, damp=damp, baselevel=bl)
    c, bn = O.lmgc_iter(H, top, p, rhs)
    want = (np.eye(n2) - E) @ np.linalg.solve(A, rhs)
    assert np.max(np.abs(c - want)) < 1e-11 * np.max(np.abs(want))
    assert np.max(np.abs(bn - (rhs - A @ c))) < 1e-13 * np.max(np.abs(rhs))
