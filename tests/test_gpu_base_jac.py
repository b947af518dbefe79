"""GPU parity of the base ls over jac (LinearSolver np/procs/ls.cc:637-749
with Smoother np/procs/iter.cc:817-842 on the base level: up to $m damped
Jacobi steps, the base stop test on each defect; ugmg_params.base_iter)
against the oracle, on the 2D Q1 Poisson problem (scalar) and the 2D P1
coupled system b = 2 with a $damp per component: ugmg_base_solve, one lmgc
cycle, and ls solves (FP64: every iteration's defects within 1e-9 relative
and the same iteration count, with a base $red that stops the base loop
early and one that never does; mixed precision to the tolerance)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module", params=["poisson_q1_b1", "p1_b2"])
def system(request):
    from ug_amd import problems
    print(f"\n[base_jac] system {request.param}", flush=True)
    P = problems.q1(2, 2, 6) if request.param == "poisson_q1_b1" else problems.p1_local(6, 0, 2)
    lv = P.levels[-1]
    rhs = np.random.default_rng(0x5EED0F01).uniform(-1, 1, lv.b * lv.n)
    rhs[np.repeat(lv.skip != 0, lv.b)] = 0.0
    P.rhs = rhs
    return P


def _bd(P):
    return [0.7, 0.6][:P.levels[-1].b]


def _ctx(P, m, red, **kw):
    from ug_amd import ugmg
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=2, smoother=ugmg.SMOOTH_GS, damp=1.0, base_damp=_bd(P),
                   base_abslimit=1e-300, base_red=red, base_maxiter=m, base_iter=ugmg.BASE_JAC, **kw)
    ctx.preprocess()
    return ctx


def _op(P, m, red):
    return O.params(smoother=O.MCGS, damp=1.0, baselevel=2, base_damp=_bd(P), base_red=red, base_maxiter=m,
                    base_iter=1)


@pytest.mark.parametrize("m,red", [(30, 1e-2), (8, 1e-300)])
def test_base_cycle_and_solve(system, m, red):
    P = system
    H = O.Hierarchy(P.levels)
    B = P.levels[-1].b
    n2 = B * P.levels[2].n
    bb = np.random.default_rng(62).uniform(-1, 1, n2)
    bb[np.repeat(P.levels[2].skip != 0, B)] = 0.0
    want = O.ls_solve(H, P.top, _op(P, m, red), P.rhs, red=1e-10, maxiter=200)
    cw, bw = O.lmgc_iter(H, P.top, _op(P, m, red), P.rhs)
    # the base ls alone: lmgc_iter with the base level as the top level
    cbw, bbw = O.lmgc_iter(H, 2, _op(P, m, red), bb)
    for graph in (0, 1):
        ctx = _ctx(P, m, red, use_graph=graph)
        try:
            cd = torch.zeros(n2, dtype=torch.float64, device="cuda")
            bd = torch.from_numpy(bb.copy()).cuda()
            ctx.base_solve(cd, bd)
            torch.cuda.synchronize()
            assert _rel(cd.cpu().numpy(), cbw) < 1e-12 and _rel(bd.cpu().numpy(), bbw) < 1e-11
            c, bn = ctx.iter_host(P.top, P.rhs)
            x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=200)
        finally:
            ctx.close()
        assert _rel(c, cw) < 1e-11 and _rel(bn, bw) < 1e-10
        assert res.converged and res.number_of_linear_iterations == want["nit"], graph
        assert np.max(np.abs(d - want["defects"]) / want["defects"]) < 1e-9
        assert _rel(x, want["x"]) < 1e-9


def test_base_mixed_solve(system):
    P = system
    B = P.levels[-1].b
    want = O.ls_solve(O.Hierarchy(P.levels), P.top, _op(P, 30, 1e-2), P.rhs, red=1e-10, maxiter=200)
    ctx = _ctx(P, 30, 1e-2, precision=1, inner_cycles=2, use_graph=1)
    try:
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=200)
    finally:
        ctx.close()
    assert res.converged
    assert np.all(np.array(res.last_defect[:B]) < 1e-10 * np.array(res.first_defect[:B]))
    assert np.max(np.abs(x - want["x"])) < 1e-8 * np.max(np.abs(want["x"]))
