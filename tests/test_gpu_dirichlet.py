"""AssembleDirichletBoundary and ClearDirichletValues (np/udm/disctools.cc:
1801-1875) on levels already set: a level whose Dirichlet rows were
perturbed gets back the generator's rows (which apply UG's rule) and b_j =
x_j on its Dirichlet components only; ClearDirichletValues zeroes exactly
those components of a vector, on levels kept in list order and in the GS
device order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("b", [1, 2])
def test_assemble_dirichlet_restores_the_rows(b):
    from ug_amd import problems, ugmg
    P = problems.p1_local(4, 0, b)
    lv = P.levels[-1]
    bb = b * b
    want = lv.vals.copy()
    rng = np.random.default_rng(b)
    bad = lv.vals.copy()
    dir_rows = np.flatnonzero(lv.skip != 0)
    for v in dir_rows:
        lo, hi = lv.row_ptr[v] * bb, lv.row_ptr[v + 1] * bb
        bad[lo:hi] = rng.uniform(-1, 1, hi - lo)
    lv.vals = bad
    x = rng.uniform(-1, 1, lv.n * b)
    rhs = rng.uniform(-1, 1, lv.n * b)
    ctx = ugmg.Context(0)
    try:
        ctx.set_hierarchy(P)
        got = rhs.copy()
        ctx.assemble_dirichlet(P.top, x, got)
        mask = np.repeat(lv.skip != 0, b) if b > 1 else lv.skip != 0
        for j in range(b):
            m = ((lv.skip >> j) & 1) == 1
            assert np.array_equal(got[j::b][m], x[j::b][m]) and np.array_equal(got[j::b][~m], rhs[j::b][~m])
        assert mask.any()
        ctx.set_params(nu1=1, nu2=1, baselevel=1)
        ctx.preprocess()
        rp, ci, va = ctx.pattern(P.top)
        assert np.array_equal(va, want)
    finally:
        ctx.close()


@pytest.mark.parametrize("smoother", [0, 1])
def test_clear_dirichlet_zeroes_the_dirichlet_components(smoother):
    from ug_amd import problems, ugmg
    P = problems.p1_local(4, 0, 2)
    lv = P.levels[-1]
    ctx = ugmg.Context(0)
    try:
        ctx.set_hierarchy(P)
        ctx.set_params(nu1=1, nu2=1, baselevel=1, smoother=smoother)
        ctx.preprocess()
        x = np.random.default_rng(3).uniform(1, 2, 2 * lv.n)
        xd = torch.from_numpy(x.copy()).cuda()
        ctx.clear_dirichlet(P.top, xd)
        torch.cuda.synchronize()
        got = xd.cpu().numpy()
        for j in range(2):
            m = ((lv.skip >> j) & 1) == 1
            assert np.all(got[j::2][m] == 0.0) and np.array_equal(got[j::2][~m], x[j::2][~m])
    finally:
        ctx.close()
