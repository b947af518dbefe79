"""Parity on UG-like connectivity orders (GPU against the oracle).

UG lists a level's vectors in creation order and each row's connections
newest first (gm/dlmgr.t:401-420, gm/algebra.cc:1042-1075); the builders
produce lexicographic lists and ascending rows.  problems.reordered gives
hierarchies whose vector lists are seeded random permutations (level by
level, the transfer mapped on both sides) and whose rows list their
off-diagonal entries newest first.  On those inputs the device follows the
arrays it is given exactly as the oracle (UG's loops restated) does: the
greedy colouring in the new list order, row sums in the new entry order,
Jacobi and multicolour-GS ls+lmgc defects per iteration within the bars of
the lexicographic tests, block (3x3) levels, and the ON_SURFACE solve of a
locally refined hierarchy.  A reordering changes nothing but rounding in a
Jacobi solve: its solution, mapped back, agrees with the original's.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    torch.cuda.synchronize()
    return t


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _ctx(P, smoother, bl, damp, **kw):
    from ug_amd import ugmg
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=bl, smoother=smoother, damp=damp,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, use_graph=1, **kw)
    ctx.preprocess()
    return ctx


@pytest.fixture(scope="module")
def jump33_r(jump3d_33):
    from ug_amd import problems
    return problems.reordered(jump3d_33, 0x5EED0201)


def _solve(ctx, P, red=1e-10, maxiter=200):
    n = P.levels[-1].n * P.levels[-1].b
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    b = _dev(P.rhs)
    res, defects = ctx.solve(P.top, x, b, red=red, maxiter=maxiter)
    return res, np.array(defects), _host(x)


def test_colouring_follows_the_list_order(jump33_r):
    from ug_amd import ugmg
    ctx = _ctx(jump33_r, ugmg.SMOOTH_GS, 1, 1.0)
    H = O.Hierarchy(jump33_r.levels)
    try:
        for l in range(2, jump33_r.top + 1):
            nc, want = O.greedy_color(H, l)
            assert np.array_equal(ctx.colors(l), want), l
            assert ctx.colors(l).max() + 1 == nc
    finally:
        ctx.close()


def test_jacobi_defects_match_oracle_on_ug_order(jump3d_33, jump33_r):
    from ug_amd import ugmg
    P = jump33_r
    want = O.ls_solve(O.Hierarchy(P.levels), P.top, O.params(smoother=O.JAC, damp=0.8, baselevel=1), P.rhs,
                      red=1e-10, maxiter=200)
    ctx = _ctx(P, ugmg.SMOOTH_JAC, 1, 0.8)
    try:
        res, d, x = _solve(ctx, P)
    finally:
        ctx.close()
    assert res.number_of_linear_iterations == want["nit"]
    assert np.max(np.abs(d[:, 0] - want["defects"][:, 0]) / want["defects"][:, 0]) < 1e-10
    # the reordering only changes rounding: mapped back, the solution is the
    # lexicographic hierarchy's
    ref = O.ls_solve(O.Hierarchy(jump3d_33.levels), jump3d_33.top,
                     O.params(smoother=O.JAC, damp=0.8, baselevel=1), jump3d_33.rhs, red=1e-10, maxiter=200)
    back = np.empty_like(x)
    back[P.perms[-1]] = x
    assert np.linalg.norm(back - ref["x"]) / np.linalg.norm(ref["x"]) < 1e-9


def test_mcgs_defects_match_oracle_on_ug_order(jump33_r):
    from ug_amd import ugmg
    P = jump33_r
    H = O.Hierarchy(P.levels)
    mc = O.ls_solve(H, P.top, O.params(smoother=O.MCGS, damp=1.0, baselevel=1), P.rhs, red=1e-10, maxiter=200)
    lex = O.ls_solve(H, P.top, O.params(smoother=O.GS, damp=1.0, baselevel=1), P.rhs, red=1e-10, maxiter=200)
    ctx = _ctx(P, ugmg.SMOOTH_GS, 1, 1.0)
    try:
        res, d, x = _solve(ctx, P)
    finally:
        ctx.close()
    assert res.converged and res.number_of_linear_iterations == mc["nit"]
    assert np.max(np.abs(d[:, 0] - mc["defects"][:, 0]) / mc["defects"][:, 0]) < 1e-9
    assert np.linalg.norm(x - lex["x"]) / np.linalg.norm(lex["x"]) < 1e-8


def test_block_levels_on_ug_order():
    """3x3 blocks (P1 elasticity): block rows reordered as whole blocks."""
    from ug_amd import problems, ugmg
    P = problems.reordered(problems.elasticity(4), 0x5EED0202)
    H = O.Hierarchy(P.levels)
    for sm, osm, damp in ((ugmg.SMOOTH_GS, O.MCGS, 1.0), (ugmg.SMOOTH_JAC, O.JAC, 0.6)):
        want = O.ls_solve(H, P.top, O.params(smoother=osm, damp=damp, baselevel=1), P.rhs, red=1e-8,
                          maxiter=200)
        ctx = _ctx(P, sm, 1, damp)
        try:
            res, d, x = _solve(ctx, P, red=1e-8)
        finally:
            ctx.close()
        assert res.number_of_linear_iterations == want["nit"], sm
        w = want["defects"][: len(d)]
        assert np.max(np.abs(d - w) / np.maximum(w, 1e-300)) < 1e-9, sm


def test_surface_solve_on_ug_order():
    """ON_SURFACE solve of a red-green hierarchy with reordered lists: the
    per-iteration surface defects and every level's x of the oracle's
    restatement."""
    from ug_amd import problems, ugmg
    P = problems.reordered(problems.p1_local(5, 3), 0x5EED0203)
    bl = 1
    assert problems.fullrefinelevel(P) < P.top
    fs = problems.surface_rhs(P, bl, 0x5EED0204)
    want = O.ls_solve_surface(O.Hierarchy(P.levels), P.top, O.params(baselevel=bl, smoother=O.JAC, damp=0.8),
                              fs, red=1e-10, maxiter=40)
    ctx = _ctx(P, ugmg.SMOOTH_JAC, bl, 0.8)
    try:
        xs, bs, res, d = ctx.solve_surface_host(P.top, fs, red=1e-10, maxiter=40)
    finally:
        ctx.close()
    assert res.number_of_linear_iterations == want["nit"]
    assert d.shape == want["defects"].shape
    assert np.all(np.abs(d - want["defects"]) <= 1e-10 * want["defects"])
    for l in range(bl, P.top + 1):
        assert np.max(np.abs(xs[l] - want["x"][l])) <= 1e-10 * np.max(np.abs(want["x"][l])), l
