"""CPU pins of the oracle's transfer over an interpolation matrix (UG's
transfer $M: RestrictByMatrix / InterpolateCorrectionByMatrix,
np/algebra/transgrid.cc:1160-1380): b x b blocks per entry,
M[i, j] = coarse component i, fine component j.

* identity blocks times the standard weights give StandardRestrict /
  StandardInterpolateCorrection (restriction bit for bit);
* the operators as dense block matrices: restriction diag(coarse VECSKIP
  clear) P_M^T over the fine vectors of VCLASS >= 2 after the reset of the
  coarse rows of VNCLASS >= 2, $damp scaling only those rows after the sum;
  interpolation $damp diag(fine VECSKIP clear) P_M;
* one lmgc cycle = the textbook V-cycle over those operators (block Jacobi
  and block MC-GS).
"""
import numpy as np
import pytest

from oracle import oracle as O
from ug_amd import problems

MIX2 = np.array([[1.0, 0.2], [-0.1, 0.9]])
MIX3 = np.array([[1.0, 0.1, -0.05], [0.05, 0.95, 0.1], [-0.1, 0.0, 1.05]])


def _dense_p(f, nc):
    """Interpolation as a dense (n_f b) x (n_c b) matrix: block (v, w) = M^T."""
    b = f.b
    Pd = np.zeros((f.n * b, nc * b))
    for v in range(f.n):
        for q in range(f.p_ptr[v], f.p_ptr[v + 1]):
            w = f.p_idx[q]
            M = f.p_blk[q * b * b:(q + 1) * b * b].reshape(b, b)
            Pd[v * b:(v + 1) * b, w * b:(w + 1) * b] += M.T
    return Pd


def _ops(P, l):
    f, c = P.levels[l], P.levels[l - 1]
    b = f.b
    Pd = _dense_p(f, c.n)
    fmask = ((f.skip[:, None] >> np.arange(b)) & 1).ravel() == 0
    cmask = ((c.skip[:, None] >> np.arange(b)) & 1).ravel() == 0
    act = np.repeat(f.vclass >= 2, b)
    R = cmask[:, None] * (Pd * act[:, None]).T
    I = fmask[:, None] * Pd
    return R, I


def test_identity_blocks_are_the_standard_transfer():
    P = problems.elasticity(3)
    Q = problems.with_imat(P, np.eye(3))
    H, HQ = O.Hierarchy(P.levels), O.Hierarchy(Q.levels)
    for l in range(1, P.top + 1):
        f, c = P.levels[l], P.levels[l - 1]
        bf = np.random.default_rng(l).uniform(-1, 1, f.n * 3)
        assert np.array_equal(O.restrict(H, l, bf), O.restrict(HQ, l, bf))
        cc = np.random.default_rng(10 + l).uniform(-1, 1, c.n * 3)
        a, b = O.interpolate(H, l, cc, damp=0.7), O.interpolate(HQ, l, cc, damp=0.7)
        assert np.max(np.abs(a - b)) <= 1e-15 * np.max(np.abs(a))


@pytest.mark.parametrize("b,mix", [(2, MIX2), (3, MIX3)])
def test_matrix_transfer_is_dense_block_operator(b, mix):
    P = problems.with_imat(problems.p1_local(3, 2, b, coupling=np.eye(b) * 2.0 + 0.1), mix)
    H = O.Hierarchy(P.levels)
    for l in range(1, P.top + 1):
        f, c = P.levels[l], P.levels[l - 1]
        R, I = _ops(P, l)
        bf = np.random.default_rng(l).uniform(-1, 1, f.n * b)
        to_c = np.full(c.n * b, 7.0)
        reset = np.repeat(c.vnclass >= 2, b)
        want = np.where(reset, 0.0, 7.0) + R @ bf
        assert np.max(np.abs(O.restrict(H, l, bf, to_c=to_c) - want)) < 1e-13
        # $damp scales the reset rows only, after the sum
        want_d = np.where(reset, 0.5 * (R @ bf), 7.0 + R @ bf)
        assert np.max(np.abs(O.restrict(H, l, bf, damp=0.5, to_c=to_c) - want_d)) < 1e-13
        cc = np.random.default_rng(10 + l).uniform(-1, 1, c.n * b)
        assert np.max(np.abs(O.interpolate(H, l, cc, damp=0.7) - 0.7 * (I @ cc))) < 1e-13


def _mg_error(P, top, bl, kind, damp, nu1, nu2):
    H = O.Hierarchy(P.levels)
    b = P.levels[0].b
    E = {bl: np.zeros((b * P.levels[bl].n,) * 2)}
    for l in range(bl + 1, top + 1):
        f, c = P.levels[l], P.levels[l - 1]
        n = f.n
        A = f.dense()
        if kind == "jac":
            blk = np.kron(np.eye(n), np.ones((b, b))).astype(bool)
        else:
            _, col = O.greedy_color(H, l)
            rk = col.astype(np.int64) * n + np.arange(n)
            blk = np.kron(rk[None, :] <= rk[:, None], np.ones((b, b))).astype(bool)
        S = np.eye(b * n) - damp * np.linalg.solve(np.where(blk, A, 0.0), A)
        R, I = _ops(P, l)
        CG = np.eye(b * n) - I @ (np.eye(b * c.n) - E[l - 1]) @ np.linalg.solve(c.dense(), R @ A)
        E[l] = np.linalg.matrix_power(S, nu2) @ CG @ np.linalg.matrix_power(S, nu1)
    return E[top]


@pytest.mark.parametrize("kind,damp", [("jac", 0.6), ("mcgs", 1.0)])
def test_lmgc_cycle_over_an_interpolation_matrix_is_textbook(kind, damp):
    P = problems.with_imat(problems.p1_local(3, 0, 2), MIX2)
    H = O.Hierarchy(P.levels)
    top, bl = P.top, 1
    E = _mg_error(P, top, bl, kind, damp, 2, 2)
    A = P.levels[top].dense()
    rhs = np.random.default_rng(4).uniform(-1, 1, A.shape[0])
    rhs[np.repeat(P.levels[top].skip != 0, 2)] = 0.0
    p = O.params(smoother=O.JAC if kind == "jac" else O.MCGS, damp=damp, baselevel=bl)
    c, bn = O.lmgc_iter(H, top, p, rhs)
    want = (np.eye(A.shape[0]) - E) @ np.linalg.solve(A, rhs)
    assert np.max(np.abs(c - want)) < 1e-11 * np.max(np.abs(want))
    assert np.max(np.abs(bn - (rhs - A @ c))) < 1e-13 * np.max(np.abs(rhs))
