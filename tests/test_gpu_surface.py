"""GPU parity of the solve on locally refined hierarchies (ugmg_solve_surface
through the C ABI) against the oracle's restatement.

Hierarchies: problems.p1_local (2D P1, uniform levels, then red-green
refinement toward one corner, UG's vector classes), scalar and 3x3 blocks.
Bars: with damped Jacobi every iteration's surface defect norm within 1e-10
relative of the oracle's and x on every level within 1e-10; with
multicolour GS (the oracle ordered by the same greedy colours) the defects
and the final solutions within 1e-8.  On a uniformly refined hierarchy the
surface solve equals ugmg_solve on the same context bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ug_amd import problems, ugmg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

BL = 1


def _ctx(Q, smoother, damp, precision=None):
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(Q)
    kw = {} if precision is None else dict(precision=precision)
    ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=BL, smoother=smoother, damp=damp,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, **kw)
    ctx.preprocess()
    return ctx


def _both(Q, dev_sm, orc_sm, damp, maxiter, red=1e-10, seed=5):
    fs = problems.surface_rhs(Q, BL, seed)
    ctx = _ctx(Q, dev_sm, damp)
    assert ctx.fullrefinelevel() == problems.fullrefinelevel(Q)
    xs, bs, res, d = ctx.solve_surface_host(Q.top, fs, red=red, maxiter=maxiter)
    ctx.close()
    H = O.Hierarchy(Q.levels)
    p = O.params(baselevel=BL, smoother=orc_sm, damp=damp)
    r = O.ls_solve_surface(H, Q.top, p, fs, red=red, maxiter=maxiter)
    return xs, bs, res, d, r


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("nc,depth", [(1, 1), (1, 3), (3, 2)])
def test_jacobi_defects_and_x_on_every_level(nc, depth):
    Q = problems.p1_local(5, depth, nc)
    assert problems.fullrefinelevel(Q) < Q.top
    print(f"surface jacobi nc={nc} depth={depth} n_top={Q.levels[-1].n}", flush=True)
    xs, bs, res, d, r = _both(Q, ugmg.SMOOTH_JAC, O.JAC, 0.8, 40)
    assert res.converged and r["converged"]
    assert res.number_of_linear_iterations == r["nit"]
    assert d.shape == r["defects"].shape
    assert np.all(np.abs(d - r["defects"]) <= 1e-10 * r["defects"])
    for l in range(BL, Q.top + 1):
        assert _rel(xs[l], r["x"][l]) <= 1e-10, l


@pytest.mark.parametrize("nc,depth", [(1, 2), (3, 3)])
def test_mcgs_solution_on_every_level(nc, depth):
    Q = problems.p1_local(5, depth, nc)
    print(f"surface mcgs nc={nc} depth={depth} n_top={Q.levels[-1].n}", flush=True)
    xs, bs, res, d, r = _both(Q, ugmg.SMOOTH_GS, O.MCGS, 1.0, 40)
    assert res.converged and r["converged"]
    assert abs(res.number_of_linear_iterations - r["nit"]) <= 1
    k = min(len(d), len(r["defects"]))
    assert np.all(np.abs(d[:k] - r["defects"][:k]) <= 1e-8 * r["defects"][:k])
    for l in range(BL, Q.top + 1):
        assert _rel(xs[l], r["x"][l]) <= 1e-8, l
    # the surface rows hold b = f - A x on their own level
    fs = problems.surface_rhs(Q, BL, 5)
    frl = problems.fullrefinelevel(Q)
    for l in range(frl, Q.top + 1):
        lv = Q.levels[l]
        m = problems.fine_grid_dof(lv) if l < Q.top else lv.vclass >= 2
        m = np.repeat(m, lv.b)
        res_l = fs[l] - lv.dense() @ xs[l]
        assert np.max(np.abs(res_l[m] - bs[l][m])) <= 1e-12


@pytest.mark.parametrize("smoother,damp", [(ugmg.SMOOTH_JAC, 0.8), (ugmg.SMOOTH_GS, 1.0)])
def test_uniform_hierarchy_equals_the_plain_solve(smoother, damp):
    Q = problems.p1_local(5, 0)
    fs = problems.surface_rhs(Q, BL, 9)
    ctx = _ctx(Q, smoother, damp)
    assert ctx.fullrefinelevel() == Q.top
    xs, bs, res, d = ctx.solve_surface_host(Q.top, fs, maxiter=30)
    x, b, res2, d2 = ctx.solve_host(Q.top, fs[Q.top], maxiter=30)
    ctx.close()
    assert res.number_of_linear_iterations == res2.number_of_linear_iterations
    assert np.array_equal(d, d2)
    assert np.array_equal(xs[Q.top], x)
    assert np.array_equal(bs[Q.top], b)


def test_negative_baselevel_searches_from_fullrefinelevel():
    """LmgcInit (np/procs/iter.cc:7643-7648): $b -N takes, from FULLREFINELEVEL
    down, the first level with at most N vectors (level 0 when none has);
    the locally refined levels above FULLREFINELEVEL are never chosen."""
    Q = problems.p1_local(4, 2)
    frl = problems.fullrefinelevel(Q)
    sizes = [lv.n for lv in Q.levels]
    assert frl < Q.top and sizes[frl - 1] < sizes[frl]
    for N, want in ((10 ** 9, frl), (sizes[frl - 1], frl - 1), (sizes[1], 1), (1, 0)):
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(Q)
        ctx.set_params(baselevel=-N, smoother=ugmg.SMOOTH_JAC, damp=0.8, base_abslimit=1e-300)
        ctx.preprocess()
        assert ctx.base_level() == want, (N, ctx.base_level(), want)
        ctx.close()


def test_refusals():
    Q = problems.p1_local(3, 2)
    frl = problems.fullrefinelevel(Q)
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(Q)
    ctx.set_params(baselevel=frl + 1, smoother=ugmg.SMOOTH_JAC, damp=0.8)
    ctx.preprocess()
    fs = problems.surface_rhs(Q, frl + 1, 1)
    with pytest.raises(ugmg.UgmgError):
        ctx.solve_surface_host(Q.top, fs, maxiter=2)      # FULLREFINELEVEL below the base level
    ctx.close()
