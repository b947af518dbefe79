"""CPU pins of the oracle on 2x2 and 4x4 blocks (SolveSmallBlock's n = 2
form and its Gauss elimination for n >= 4, np/algebra/block.cc:119-199):
the 2D P1 coupled system of problems.p1_local(.., ncomp=b) is the scalar
stiffness times an SPD b x b coupling, and one lmgc cycle with block Jacobi
or block MC-GS equals the textbook V-cycle operator built densely with
kron(P, I_b) transfers masked per Dirichlet component; a block Jacobi step
whose blocks need the row exchange equals the dense block solves.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ug_amd import problems


def _ops(P, l, kind, damp, H):
    """A and the smoother's error operator I - D M^-1 A, D the damping per
    component (a scalar: the same for all)."""
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
    D = np.tile(np.broadcast_to(np.asarray(damp, float), (b,)), n)
    return A, np.eye(b * n) - D[:, None] * np.linalg.solve(M, A)


def mg_error(P, top, bl, kind, damp, nu1, nu2, lmgc_damp=1.0, base_damp=1.0, post=None, post_damp=None,
             base_jac=0):
    """Error operator of one V-cycle: smoothers with $damp, the coarse
    correction interpolated and scaled by lmgc's $damp (dscalx), the base
    level solved exactly and scaled by ex's $damp, each per component; post,
    post_damp: the post-smoother's kind and $damp (None: the pre-smoother's);
    base_jac = m > 0: the base ls iterates m Jacobi steps with ex's $damp
    (no stop test met) instead of the exact solve."""
    H = O.Hierarchy(P.levels)
    b = P.levels[0].b
    nbl = P.levels[bl].n
    E = {bl: np.eye(b * nbl) - np.diag(np.tile(np.broadcast_to(np.asarray(base_damp, float), (b,)), nbl))}
    if base_jac:
        E[bl] = np.linalg.matrix_power(_ops(P, bl, "jac", base_damp, H)[1], base_jac)
    for l in range(bl + 1, top + 1):
        f, c = P.levels[l], P.levels[l - 1]
        A, S = _ops(P, l, kind, damp, H)
        _, S2 = _ops(P, l, post or kind, damp if post_damp is None else post_damp, H)
        Ac = c.dense()
        Pb = np.kron(f.dense_p(c.n), np.eye(b))
        fmask = ((f.skip[:, None] >> np.arange(b)) & 1).ravel() == 0
        cmask = ((c.skip[:, None] >> np.arange(b)) & 1).ravel() == 0
        Dl = np.tile(np.broadcast_to(np.asarray(lmgc_damp, float), (b,)), f.n)
        CG = np.eye(A.shape[0]) - (Dl * fmask)[:, None] * Pb @ (np.eye(Ac.shape[0]) - E[l - 1]) @ \
            np.linalg.solve(Ac, (Pb * cmask[None, :]).T @ A)
        E[l] = np.linalg.matrix_power(S2, nu2) @ CG @ np.linalg.matrix_power(S, nu1)
    return E[top]


COUPLING = {2: np.array([[2.0, 0.5], [0.5, 2.0]]),
            4: np.array([[2.0, 0.5, 0.2, 0.1], [0.5, 2.0, 0.3, 0.15], [0.2, 0.3, 2.0, 0.25],
                         [0.1, 0.15, 0.25, 2.0]])}
# (0, 0) far below 1e-10 times any stiffness diagonal: every interior block
# takes SolveSmallBlock's row exchange
PIVOTING = np.array([[1e-12, 1.0, 0.2, 0.1], [1.0, 2.0, 0.3, 0.1], [0.2, 0.3, 2.0, 0.25], [0.1, 0.1, 0.25, 2.0]])


@pytest.mark.parametrize("b", [2, 4])
def test_coupled_system_is_stiffness_times_coupling(b):
    P = problems.p1_local(2, 0, b)
    s = problems.p1_local(2, 0, 1)
    for a, sl in zip(P.levels, s.levels):
        assert a.b == b
        K = np.kron(sl.dense(), COUPLING[b])
        free = np.repeat(sl.skip == 0, b)
        assert np.max(np.abs(a.dense()[free] - K[free])) < 1e-14
        assert np.array_equal(a.skip, np.where(sl.skip != 0, (1 << b) - 1, 0))


@pytest.mark.parametrize("b", [2, 4])
@pytest.mark.parametrize("kind,damp", [("jac", 0.6), ("mcgs", 1.0)])
def test_block_lmgc_cycle_is_textbook_vcycle(b, kind, damp):
    P = problems.p1_local(3, 0, b)
    H = O.Hierarchy(P.levels)
    top, bl = P.top, 1
    E = mg_error(P, top, bl, kind, damp, 2, 2)
    A = P.levels[top].dense()
    n2 = A.shape[0]
    rhs = np.random.default_rng(4).uniform(-1, 1, n2)
    rhs[np.repeat(P.levels[top].skip != 0, b)] = 0.0
    p = O.params(smoother=O.JAC if kind == "jac" else O.MCGS, damp=damp, baselevel=bl)
    c, bn = O.lmgc_iter(H, top, p, rhs)
    want = (np.eye(n2) - E) @ np.linalg.solve(A, rhs)
    assert np.max(np.abs(c - want)) < 1e-11 * np.max(np.abs(want))
    assert np.max(np.abs(bn - (rhs - A @ c))) < 1e-13 * np.max(np.abs(rhs))


@pytest.mark.parametrize("b", [2, 4])
@pytest.mark.parametrize("kind", ["jac", "mcgs"])
def test_block_lmgc_cycle_with_damping_per_component(b, kind):
    """Smoother $damp, lmgc $damp and ex $damp with a factor per component
    (dscalx, np/algebra/ugblas.cc:2481-2515: Smoother np/procs/iter.cc:836,
    Lmgc's InterpolateCorrection with its $damp iter.cc:7886-7887, EXSmoother
    iter.cc:9055) against the textbook V-cycle with D = diag(damp) in each
    place."""
    P = problems.p1_local(3, 0, b)
    H = O.Hierarchy(P.levels)
    top, bl = P.top, 1
    dm, dl, db = [0.5, 0.7, 0.6, 0.8][:b], [0.9, 0.75, 1.0, 0.85][:b], [0.95, 0.8, 0.9, 0.7][:b]
    E = mg_error(P, top, bl, kind, dm, 2, 2, lmgc_damp=dl, base_damp=db)
    A = P.levels[top].dense()
    n2 = A.shape[0]
    rhs = np.random.default_rng(5).uniform(-1, 1, n2)
    rhs[np.repeat(P.levels[top].skip != 0, b)] = 0.0
    p = O.params(smoother=O.JAC if kind == "jac" else O.MCGS, damp=dm, lmgc_damp=dl, base_damp=db,
                 baselevel=bl)
    c, bn = O.lmgc_iter(H, top, p, rhs)
    want = (np.eye(n2) - E) @ np.linalg.solve(A, rhs)
    assert np.max(np.abs(c - want)) < 1e-11 * np.max(np.abs(want))
    assert np.max(np.abs(bn - (rhs - A @ c))) < 1e-13 * np.max(np.abs(rhs))
    # a scalar factor is the same as that factor for every component
    q = O.params(smoother=O.JAC if kind == "jac" else O.MCGS, damp=[0.6] * b, baselevel=bl)
    c1, _ = O.lmgc_iter(H, top, q, rhs)
    c2, _ = O.lmgc_iter(H, top, O.params(smoother=q.smoother, damp=0.6, baselevel=bl), rhs)
    assert np.array_equal(c1, c2)


@pytest.mark.parametrize("b", [1, 2])
@pytest.mark.parametrize("pre,post", [("jac", "mcgs"), ("mcgs", "jac"), ("mcgs", "mcgs")])
def test_lmgc_cycle_with_distinct_pre_and_post_smoothers(b, pre, post):
    """Lmgc's $S pre post base with two smoother objects (PreSmooth nu1
    times, np/procs/iter.cc:7812; PostSmooth nu2 times, iter.cc:7922): E = S_post^nu2 CG S_pre^nu1
    with each smoother's class and $damp."""
    P = problems.p1_local(3, 0, b)
    H = O.Hierarchy(P.levels)
    top, bl = P.top, 1
    dpre, dpost = [0.6, 0.7][:b], [0.9, 0.8][:b]
    E = mg_error(P, top, bl, pre, dpre, 2, 1, post=post, post_damp=dpost)
    A = P.levels[top].dense()
    n2 = A.shape[0]
    rhs = np.random.default_rng(6).uniform(-1, 1, n2)
    rhs[np.repeat(P.levels[top].skip != 0, b)] = 0.0
    kind = {"jac": O.JAC, "mcgs": O.MCGS}
    p = O.params(smoother=kind[pre], damp=dpre, nu1=2, nu2=1, baselevel=bl, post=kind[post], post_damp=dpost)
    c, bn = O.lmgc_iter(H, top, p, rhs)
    want = (np.eye(n2) - E) @ np.linalg.solve(A, rhs)
    assert np.max(np.abs(c - want)) < 1e-11 * np.max(np.abs(want))
    assert np.max(np.abs(bn - (rhs - A @ c))) < 1e-13 * np.max(np.abs(rhs))


@pytest.mark.parametrize("b", [1, 2])
def test_lmgc_cycle_with_jacobi_base_solver(b):
    """The base ls over jac (LinearSolver np/procs/ls.cc:637-749 with
    Smoother iter.cc:817-842 on the base level): m steps with no stop test met give
    E_base = S_jac^m; a base $abslimit met at once leaves the coarse
    correction zero (E_base = I); a $red met after k steps equals m = k."""
    P = problems.p1_local(3, 0, b)
    H = O.Hierarchy(P.levels)
    top, bl = P.top, 1
    db = [0.8, 0.7][:b]
    A = P.levels[top].dense()
    n2 = A.shape[0]
    rhs = np.random.default_rng(8).uniform(-1, 1, n2)
    rhs[np.repeat(P.levels[top].skip != 0, b)] = 0.0
    for m in (1, 4):
        E = mg_error(P, top, bl, "jac", 0.6, 2, 2, base_damp=db, base_jac=m)
        p = O.params(smoother=O.JAC, damp=0.6, base_damp=db, baselevel=bl, base_maxiter=m, base_iter=1)
        c, _ = O.lmgc_iter(H, top, p, rhs)
        want = (np.eye(n2) - E) @ np.linalg.solve(A, rhs)
        assert np.max(np.abs(c - want)) < 1e-11 * np.max(np.abs(want))
    E = mg_error(P, top, bl, "jac", 0.6, 2, 2, base_jac=0)
    E_none = mg_error(P, top, bl, "jac", 0.6, 2, 2, base_damp=0.0)    # coarse correction 0
    c, _ = O.lmgc_iter(H, top, O.params(smoother=O.JAC, damp=0.6, baselevel=bl, base_maxiter=4, base_iter=1,
                                         base_abslimit=1e300), rhs)
    want = (np.eye(n2) - E_none) @ np.linalg.solve(A, rhs)
    assert np.max(np.abs(c - want)) < 1e-11 * np.max(np.abs(want))
    # the base defect after k Jacobi steps, computed directly: the first k
    # whose norm (every component) is below 0.5 of the first stops the loop
    pk = O.params(smoother=O.JAC, damp=0.6, baselevel=bl, base_maxiter=40, base_iter=1, base_red=0.5)
    c40, _ = O.lmgc_iter(H, top, pk, rhs)
    outs = [O.lmgc_iter(H, top, O.params(smoother=O.JAC, damp=0.6, baselevel=bl, base_maxiter=k, base_iter=1,
                                         base_red=1e-300), rhs)[0] for k in range(1, 41)]
    hits = [k for k in range(1, 41) if np.array_equal(outs[k - 1], c40)]
    assert hits and hits[0] < 40
    assert not np.array_equal(outs[hits[0] - 2] if hits[0] > 1 else np.zeros_like(c40), c40)


@pytest.mark.parametrize("b", [1, 2])
def test_jacobi_base_solver_stops_where_its_defect_meets_red(b):
    """The base ls over jac alone (lmgc on the base level): numpy Jacobi
    steps x_k = damp M^-1 b_k (zero on rows of VCLASS < 3), b_k+1 = b_k - A x_k,
    stopping at the first k whose per-component defect norms are all below
    $red times the first; c = the sum of the steps taken."""
    P = problems.p1_local(3, 0, b)
    H = O.Hierarchy(P.levels)
    bl = 1
    lv = P.levels[bl]
    A = lv.dense()
    n2 = A.shape[0]
    D = np.kron(np.eye(lv.n), np.ones((b, b))).astype(bool)
    M = np.where(D, A, 0.0)
    act = np.repeat(lv.vclass >= 3, b)
    damp = np.tile(np.array([0.8, 0.7][:b]), lv.n)
    rhs = np.random.default_rng(9).uniform(-1, 1, n2)
    rhs[np.repeat(lv.skip != 0, b)] = 0.0
    for red in (0.5, 0.05, 1e-3):
        r, c = rhs.copy(), np.zeros(n2)
        first = np.linalg.norm(r.reshape(-1, b), axis=0)
        for k in range(60):
            x = np.where(act, damp * np.linalg.solve(M, r), 0.0)
            c += x
            r = r - A @ x
            if np.all(np.linalg.norm(r.reshape(-1, b), axis=0) < red * first):
                break
        p = O.params(smoother=O.JAC, damp=0.6, baselevel=bl, base_damp=[0.8, 0.7][:b], base_maxiter=60,
                     base_iter=1, base_red=red)
        cw, bw = O.lmgc_iter(H, bl, p, rhs)
        assert k < 59
        assert np.max(np.abs(cw - c)) < 1e-12 * np.max(np.abs(c)), red
        assert np.max(np.abs(bw - r)) < 1e-12 * np.max(np.abs(rhs)), red


def test_block4_jacobi_step_with_row_exchange_is_dense_block_solve():
    P = problems.p1_local(2, 0, 4, coupling=PIVOTING)
    H = O.Hierarchy(P.levels)
    lv = P.levels[-1]
    A = lv.dense()
    d = np.random.default_rng(7).uniform(-1, 1, 4 * lv.n)
    rc, x = O.l_jac(H, P.top, d)
    assert rc == 0
    for i in range(lv.n):
        s = slice(4 * i, 4 * i + 4)
        want = np.zeros(4) if lv.vclass[i] < 3 else np.linalg.solve(A[s, s], d[s])
        assert np.max(np.abs(x[s] - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))
    inner = np.flatnonzero((lv.skip == 0) & (lv.vclass >= 3))
    assert len(inner) > 0 and np.all(np.abs(A[4 * inner, 4 * inner]) < 1e-10)   # the exchange taken


def test_block4_singular_diagonal_block_is_refused():
    P = problems.p1_local(2, 0, 4, coupling=PIVOTING)
    lv = P.levels[-1]
    i = int(np.flatnonzero((lv.skip == 0) & (lv.vclass >= 3))[0])
    blk = lv.vals[lv.row_ptr[i] * 16:(lv.row_ptr[i] + 1) * 16].reshape(4, 4)
    blk[:, 0] = [1e-26, 1e-26, 0.0, 0.0]              # column 0 below SMALL_D * 1e-10 = 2.2e-25
    rc, _ = O.l_jac(O.Hierarchy(P.levels), P.top, np.ones(4 * lv.n))
    assert rc == 1
