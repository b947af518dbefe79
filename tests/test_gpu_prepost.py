"""GPU parity of lmgc with two different smoother objects ($S pre post base:
PreSmooth nu1 times, np/procs/iter.cc:7812; PostSmooth nu2 times,
iter.cc:7922) against the oracle: pre Jacobi / post MC-GS and the reverse,
with their own $damp, on the 2D Q1 Poisson problem (scalar: the folded
colour-0 sweeps and fused GS steps on the GS side) and the 2D P1 coupled
system b = 2 (block kernels); FP64 solves every iteration's defects within
1e-9 relative and the same iteration count, the mixed form to the stopping
tolerance; one cycle within 1e-11."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASES = [("jac", 0.7, "gs", 1.0), ("gs", 1.0, "jac", 0.6), ("gs", 1.0, "gs", 0.8)]


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module", params=["poisson_q1_b1", "p1_b2"])
def system(request):
    from ug_amd import problems
    print(f"\n[prepost] system {request.param}", flush=True)
    P = problems.q1(2, 2, 6) if request.param == "poisson_q1_b1" else problems.p1_local(6, 0, 2)
    lv = P.levels[-1]
    rhs = np.random.default_rng(0x5EED0E01).uniform(-1, 1, lv.b * lv.n)
    rhs[np.repeat(lv.skip != 0, lv.b)] = 0.0
    P.rhs = rhs
    return P


def _ctx(P, case, **kw):
    from ug_amd import ugmg
    pre, dpre, post, dpost = case
    k = {"jac": ugmg.SMOOTH_JAC, "gs": ugmg.SMOOTH_GS}
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=2, smoother=k[pre], damp=dpre, post_set=1,
                   post_smoother=k[post], post_damp=dpost, base_abslimit=1e-300, base_red=1e-12,
                   base_maxiter=1, **kw)
    ctx.preprocess()
    return ctx


def _oparams(case):
    pre, dpre, post, dpost = case
    k = {"jac": O.JAC, "gs": O.MCGS}
    return O.params(smoother=k[pre], damp=dpre, baselevel=2, post=k[post], post_damp=dpost)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}{c[1]}-{c[2]}{c[3]}")
def test_cycle_and_fp64_solve(system, case):
    P = system
    H = O.Hierarchy(P.levels)
    want = O.ls_solve(H, P.top, _oparams(case), P.rhs, red=1e-10, maxiter=200)
    cw, bw = O.lmgc_iter(H, P.top, _oparams(case), P.rhs)
    for graph in (0, 1):
        ctx = _ctx(P, case, use_graph=graph)
        try:
            c, bn = ctx.iter_host(P.top, P.rhs)
            x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=200)
        finally:
            ctx.close()
        assert _rel(c, cw) < 1e-11 and _rel(bn, bw) < 1e-10
        assert res.converged and res.number_of_linear_iterations == want["nit"], graph
        assert np.max(np.abs(d - want["defects"]) / want["defects"]) < 1e-9
        assert _rel(x, want["x"]) < 1e-9


@pytest.mark.parametrize("case", CASES[:2], ids=lambda c: f"{c[0]}{c[1]}-{c[2]}{c[3]}")
def test_mixed_solve(system, case):
    P = system
    B = P.levels[-1].b
    want = O.ls_solve(O.Hierarchy(P.levels), P.top, _oparams(case), P.rhs, red=1e-10, maxiter=200)
    ctx = _ctx(P, case, precision=1, inner_cycles=2, use_graph=1)
    try:
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=200)
    finally:
        ctx.close()
    assert res.converged
    assert np.all(np.array(res.last_defect[:B]) < 1e-10 * np.array(res.first_defect[:B]))
    assert np.max(np.abs(x - want["x"])) < 1e-8 * np.max(np.abs(want["x"]))
