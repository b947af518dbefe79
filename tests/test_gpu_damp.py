"""GPU parity of damping factors per component against the oracle: the
smoothers', lmgc's and ex's $damp read as VEC_SCALARs and applied by dscalx
(np/algebra/ugblas.cc:2481-2515; Smoother np/procs/iter.cc:836, Lmgc's
InterpolateCorrection iter.cc:7886-7887, EXSmoother iter.cc:9055) on the 2D
P1 coupled systems b = 2 and 4 (problems.p1_local) and on 3D elasticity
with an interpolation matrix (transfer $M).  Checked: Jacobi and MC-GS
steps, restriction and interpolation through the C ABI (the standard
transfer damps each fine term, the ByMatrix transfer scales the coarse rows
of VNCLASS >= 2 after the sum, transgrid.cc:1160-1380), the ex base solve,
one lmgc cycle, and ls solves with Jacobi (every iteration's defects within
1e-10) and MC-GS (FP64 against the colour-ordered oracle; mixed precision
to the stopping tolerance).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DM = [0.55, 0.7, 0.6, 0.8]      # smoother $damp per component
DL = [0.9, 0.75, 1.0, 0.85]     # lmgc $damp
DB = [0.95, 0.8, 0.9, 0.7]      # ex $damp
MIX3 = np.eye(3) + 0.03 * np.array([[0.0, 1.0, -0.5], [0.5, -0.5, 1.0], [-1.0, 0.0, 0.5]])


def _dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    torch.cuda.synchronize()
    return t


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _rhs(P, seed):
    lv = P.levels[-1]
    rhs = np.random.default_rng(seed).uniform(-1, 1, lv.b * lv.n)
    rhs[np.repeat(lv.skip != 0, lv.b)] = 0.0
    return rhs


def _ctx(P, smoother, b, baselevel=2, **kw):
    from ug_amd import ugmg
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=baselevel, smoother=smoother, damp=DM[:b],
                   lmgc_damp=DL[:b], base_damp=DB[:b], base_abslimit=1e-300, base_red=1e-12,
                   base_maxiter=1, **kw)
    ctx.preprocess()
    return ctx


def _oparams(kind, b, bl=2):
    return O.params(smoother=kind, damp=DM[:b], lmgc_damp=DL[:b], base_damp=DB[:b], baselevel=bl)


@pytest.fixture(scope="module", params=["p1_b2", "p1_b4", "elast_b3_imat"])
def system(request):
    from ug_amd import problems
    name = request.param
    print(f"\n[damp] system {name}", flush=True)
    if name == "p1_b2":
        P = problems.p1_local(6, 0, 2)
    elif name == "p1_b4":
        P = problems.p1_local(5, 0, 4)
    else:
        P = problems.with_imat(problems.elasticity(5), MIX3)
    P.rhs = _rhs(P, 0x5EED0D01)
    return P


def test_steps_transfer_base_and_cycle(system):
    from ug_amd import ugmg
    P = system
    H = O.Hierarchy(P.levels)
    B = P.levels[-1].b
    cg, cj = _ctx(P, ugmg.SMOOTH_GS, B), _ctx(P, ugmg.SMOOTH_JAC, B)
    try:
        for l in range(3, P.top + 1):
            lv = P.levels[l]
            n2 = B * lv.n
            r = np.random.default_rng(100 + l)
            b = r.uniform(-1, 1, n2)
            xw, bw = O.smoother(H, l, O.JAC, DM[:B], b)
            xd, bd = _dev(np.zeros(n2)), _dev(b)
            cj.jacobi_step(l, xd, bd)
            assert _rel(_host(xd), xw) < 1e-13 and _rel(_host(bd), bw) < 1e-12, l
            _, col = O.greedy_color(H, l)
            rank = (col.astype(np.int64) * lv.n + np.arange(lv.n)).astype(np.int32)
            xw, bw = O.smoother(H, l, O.GS, DM[:B], b, rank=rank)
            xd, bd = _dev(np.zeros(n2)), _dev(b)
            cg.gs_step(l, xd, bd)
            assert _rel(_host(xd), xw) < 1e-12 and _rel(_host(bd), bw) < 1e-11, l
            nc2 = B * P.levels[l - 1].n
            bf, cc = r.uniform(-1, 1, n2), r.uniform(-1, 1, nc2)
            to_c = _dev(np.full(nc2, 5.0))
            cg.restrict(l, to_c, _dev(bf), damp=DL[:B])
            assert _rel(_host(to_c), O.restrict(H, l, bf, damp=DL[:B], to_c=np.full(nc2, 5.0))) < 1e-14, l
            to_f = _dev(np.full(n2, 3.0))
            cg.interpolate(l, to_f, _dev(cc), damp=DB[:B])
            assert _rel(_host(to_f), O.interpolate(H, l, cc, damp=DB[:B])) < 1e-15, l
        n2 = B * P.levels[2].n
        b = np.random.default_rng(61).uniform(-1, 1, n2)
        rc, xw, bw = O.Ex(H, 2, reorder=1).smooth(b, damp=DB[:B])
        assert rc == 0
        cd, bd = _dev(np.zeros(n2)), _dev(b)
        cg.base_solve(cd, bd)
        assert _rel(_host(cd), xw) < 1e-12
        for ctx, kind in ((cg, O.MCGS), (cj, O.JAC)):
            cw, bw = O.lmgc_iter(H, P.top, _oparams(kind, B), P.rhs)
            c, bn = ctx.iter_host(P.top, P.rhs)
            assert _rel(c, cw) < 1e-11 and _rel(bn, bw) < 1e-10, kind
    finally:
        cg.close()
        cj.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_jacobi_solve_defects(system, graph):
    from ug_amd import ugmg
    P = system
    B = P.levels[-1].b
    want = O.ls_solve(O.Hierarchy(P.levels), P.top, _oparams(O.JAC, B), P.rhs, red=1e-10, maxiter=200)
    ctx = _ctx(P, ugmg.SMOOTH_JAC, B, use_graph=graph)
    try:
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=200)
    finally:
        ctx.close()
    assert res.converged and res.number_of_linear_iterations == want["nit"]
    assert np.max(np.abs(d - want["defects"]) / want["defects"]) < 1e-10
    assert _rel(x, want["x"]) < 1e-10


@pytest.mark.parametrize("precision", [0, 1])
def test_gs_solve(system, precision):
    from ug_amd import ugmg
    P = system
    B = P.levels[-1].b
    H = O.Hierarchy(P.levels)
    mc = O.ls_solve(H, P.top, _oparams(O.MCGS, B), P.rhs, red=1e-10, maxiter=200)
    ctx = _ctx(P, ugmg.SMOOTH_GS, B, precision=precision, inner_cycles=2, use_graph=1)
    try:
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=200)
    finally:
        ctx.close()
    assert res.converged and mc["nit"] < 200
    assert np.all(np.array(res.last_defect[:B]) < 1e-10 * np.array(res.first_defect[:B]))
    assert np.max(np.abs(x - mc["x"])) < 1e-8 * np.max(np.abs(mc["x"]))
    if precision == 0:
        assert res.number_of_linear_iterations == mc["nit"]
        assert np.max(np.abs(d - mc["defects"]) / mc["defects"]) < 1e-9


def test_scalar_damp_unchanged_by_the_per_component_path(system):
    """Equal factors take the fused single-factor kernels; the same factor
    given per component through the other path agrees to rounding."""
    from ug_amd import ugmg
    P = system
    B = P.levels[-1].b
    outs = []
    for dm in (0.6, [0.6] * (B - 1) + [0.6 * (1 + 2 ** -40)]):
        ctx = ugmg.Context(0)
        try:
            ctx.set_hierarchy(P)
            ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=2, smoother=ugmg.SMOOTH_GS, damp=dm,
                           base_abslimit=1e-300, base_red=1e-12, base_maxiter=1)
            ctx.preprocess()
            outs.append(ctx.iter_host(P.top, P.rhs)[0])
        finally:
            ctx.close()
    assert _rel(outs[1], outs[0]) < 1e-10
