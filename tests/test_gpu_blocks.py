"""GPU parity of the 3x3-block path (P1 linear elasticity, b = 3) against
the oracle: block SpMV, SolveSmallBlock Jacobi and multicolour GS steps,
per-component transfer, the ex base solve over scalar unknowns, one lmgc
cycle, and ls + lmgc solves with per-component defect norms (Jacobi:
every iteration within 1e-10 relative; multicolour GS: final solution
within 1e-8), FP64 and FP32 cycle.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()
    return t


def _host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def el4():
    from ug_amd import problems
    return problems.elasticity(4, seed=21)      # 17^3 nodes, 14739 unknowns


def _ctx(P, smoother, damp, baselevel=1, **kw):
    from ug_amd import ugmg
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=baselevel, smoother=smoother, damp=damp,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, **kw)
    ctx.preprocess()
    return ctx


@pytest.fixture(scope="module")
def ctx_gs(el4):
    from ug_amd import ugmg
    c = _ctx(el4, ugmg.SMOOTH_GS, 1.0)
    yield c
    c.close()


def test_block_pattern_and_spmv(el4, ctx_gs):
    H = O.Hierarchy(el4.levels)
    for l in range(1, el4.top + 1):
        lv = el4.levels[l]
        rp, ci, va = ctx_gs.pattern(l)
        assert np.array_equal(rp, lv.row_ptr) and np.array_equal(ci, lv.col_idx)
        assert np.array_equal(va, lv.vals)
        r = np.random.default_rng(l)
        x, y = r.uniform(-1, 1, 3 * lv.n), r.uniform(-1, 1, 3 * lv.n)
        xd = _dev(x)
        ctx_gs.dmatmul_minus(l, xd, _dev(y))
        assert _rel(_host(xd), O.dmatmul_minus(H, l, x, y)) < 1e-14, l
        nrm = ctx_gs.dnrm2x(l, _dev(x))
        np.testing.assert_allclose(nrm, O.dnrm2x(H, l, x), rtol=1e-13)


def test_block_smoother_steps(el4, ctx_gs):
    from ug_amd import ugmg
    H = O.Hierarchy(el4.levels)
    cj = _ctx(el4, ugmg.SMOOTH_JAC, 0.6)
    for l in range(2, el4.top + 1):
        n3 = 3 * el4.levels[l].n
        b = np.random.default_rng(30 + l).uniform(-1, 1, n3)
        xw, bw = O.smoother(H, l, O.JAC, 0.6, b)
        xd, bd = _dev(np.zeros(n3)), _dev(b)
        cj.jacobi_step(l, xd, bd)
        assert _rel(_host(xd), xw) < 1e-13 and _rel(_host(bd), bw) < 1e-12, l
        _, col = O.greedy_color(H, l)
        rank = (col.astype(np.int64) * el4.levels[l].n + np.arange(el4.levels[l].n)).astype(np.int32)
        xw, bw = O.smoother(H, l, O.GS, 1.0, b, rank=rank)
        xd, bd = _dev(np.zeros(n3)), _dev(b)
        ctx_gs.gs_step(l, xd, bd)
        assert _rel(_host(xd), xw) < 1e-12 and _rel(_host(bd), bw) < 1e-11, l
    cj.close()


def test_block_transfer_and_base(el4, ctx_gs):
    H = O.Hierarchy(el4.levels)
    for l in range(2, el4.top + 1):
        nf, nc = 3 * el4.levels[l].n, 3 * el4.levels[l - 1].n
        bf = np.random.default_rng(40 + l).uniform(-1, 1, nf)
        cc = np.random.default_rng(50 + l).uniform(-1, 1, nc)
        to_c = _dev(np.full(nc, 5.0))
        ctx_gs.restrict(l, to_c, _dev(bf))
        assert _rel(_host(to_c), O.restrict(H, l, bf)) < 1e-14, l
        to_f = _dev(np.full(nf, 3.0))
        ctx_gs.interpolate(l, to_f, _dev(cc), damp=0.7)
        assert _rel(_host(to_f), O.interpolate(H, l, cc, damp=0.7)) < 1e-15, l
    n3 = 3 * el4.levels[1].n
    b = np.random.default_rng(60).uniform(-1, 1, n3)
    rc, xw, bw = O.Ex(H, 1, reorder=1).smooth(b)
    assert rc == 0
    cd, bd = _dev(np.zeros(n3)), _dev(b)
    ctx_gs.base_solve(cd, bd)
    assert _rel(_host(cd), xw) < 1e-12
    assert np.max(np.abs(_host(bd) - bw)) < 1e-12 * np.max(np.abs(b))


def test_block_lmgc_cycle(el4, ctx_gs):
    H = O.Hierarchy(el4.levels)
    cw, bw = O.lmgc_iter(H, el4.top, O.params(smoother=O.MCGS, damp=1.0, baselevel=1), el4.rhs)
    c, b = ctx_gs.iter_host(el4.top, el4.rhs)
    assert _rel(c, cw) < 1e-11 and _rel(b, bw) < 1e-10


def test_block_jacobi_solve_defects_per_component(el4):
    from ug_amd import ugmg
    H = O.Hierarchy(el4.levels)
    want = O.ls_solve(H, el4.top, O.params(smoother=O.JAC, damp=0.6, baselevel=1), el4.rhs, red=1e-10,
                      maxiter=200)
    for graph in (0, 1):
        ctx = _ctx(el4, ugmg.SMOOTH_JAC, 0.6, use_graph=graph)
        x, b, res, d = ctx.solve_host(el4.top, el4.rhs, red=1e-10, maxiter=200)
        assert res.number_of_linear_iterations == want["nit"] and res.converged
        assert d.shape[1] == 3
        assert np.max(np.abs(d - want["defects"]) / want["defects"]) < 1e-10
        ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_block_gs_solve_final_solution(el4, precision):
    from ug_amd import ugmg
    H = O.Hierarchy(el4.levels)
    want = O.ls_solve(H, el4.top, O.params(smoother=O.GS, damp=1.0, baselevel=1), el4.rhs, red=1e-10,
                      maxiter=200)
    ctx = _ctx(el4, ugmg.SMOOTH_GS, 1.0, precision=precision, inner_cycles=2, use_graph=1)
    x, b, res, d = ctx.solve_host(el4.top, el4.rhs, red=1e-10, maxiter=200)
    assert res.converged
    assert np.all(np.array(res.last_defect[:3]) < 1e-10 * np.array(res.first_defect[:3]))
    assert np.max(np.abs(x - want["x"])) < 1e-8 * np.max(np.abs(want["x"]))
    A = el4.levels[-1]
    ctx.close()
