"""CPU pins of problems.reordered (UG-like list orders): every level is the
same operator in new coordinates (A' = Pi A Pi^T, P' = Pi_f P Pi_c^T by
dense comparison), rows keep the diagonal first and list their off-diagonal
entries newest first (the reverse of the builder's ascending order), the
classes and Dirichlet bits move with their vectors, and the oracle's Jacobi
solve on the reordered hierarchy is the original one in new coordinates
(equal up to the rounding of the reordered sums) while its greedy colouring
is a different one.
"""
import numpy as np

from oracle import oracle as O
from ug_amd import problems


def test_reordered_levels_are_permuted_operators():
    P = problems.jump3d(4)
    Q = problems.reordered(P, 11)
    for l in range(P.top + 1):
        a, q = P.levels[l], Q.levels[l]
        pm = Q.perms[l]
        assert sorted(pm.tolist()) == list(range(a.n)) and not np.array_equal(pm, np.arange(a.n)) or a.n == 1
        if a.n <= 729:
            assert np.array_equal(q.dense(), a.dense()[np.ix_(pm, pm)])
        assert np.array_equal(q.skip, a.skip[pm]) and np.array_equal(q.vclass, a.vclass[pm])
        for r in range(0, q.n, max(1, q.n // 50)):
            ent = q.col_idx[q.row_ptr[r]:q.row_ptr[r + 1]]
            assert ent[0] == r
            orig = pm[ent[1:]]
            assert np.all(np.diff(orig) < 0)       # newest first: descending original columns
        if l > 0 and a.n <= 729:
            c = Q.perms[l - 1]
            assert np.array_equal(q.dense_p(Q.levels[l - 1].n), a.dense_p(P.levels[l - 1].n)[np.ix_(pm, c)])
    assert np.array_equal(Q.rhs, P.rhs[Q.perms[-1]])


def test_oracle_solve_is_invariant_and_colouring_is_not():
    P = problems.jump3d(4)
    Q = problems.reordered(P, 12)
    p = O.params(smoother=O.JAC, damp=0.8, baselevel=1)
    a = O.ls_solve(O.Hierarchy(P.levels), P.top, p, P.rhs, red=1e-10, maxiter=100)
    b = O.ls_solve(O.Hierarchy(Q.levels), Q.top, p, Q.rhs, red=1e-10, maxiter=100)
    assert a["nit"] == b["nit"]
    assert np.max(np.abs(a["defects"] - b["defects"]) / a["defects"]) < 1e-12
    assert np.max(np.abs(b["x"] - a["x"][Q.perms[-1]])) < 1e-12 * np.max(np.abs(a["x"]))
    _, ca = O.greedy_color(O.Hierarchy(P.levels), P.top)
    _, cb = O.greedy_color(O.Hierarchy(Q.levels), Q.top)
    assert not np.array_equal(cb, ca[Q.perms[-1]])
