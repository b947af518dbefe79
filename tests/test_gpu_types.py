"""GPU parity of hierarchies with two vector types (node and edge vectors,
b components each; np/udm/udm.h:105-108) against the oracle, on the 2D P2
system of problems.p2_uniform: per-(type, component) norms and dot products
(dnrm2x, ddot: a[aoff[VTYPE(v)] + i], ugblas.cc:3104-3176), the edge loops
of StandardRestrictNodeVector / StandardIntCorNodeVector (transgrid.cc:
188-212, 317-336), Jacobi and multicolour-GS steps, one lmgc cycle, ls +
lmgc solves with per-type defect norms (Jacobi every iteration within 1e-10
relative; MC-GS defects within 1e-9 of the colour-ordered oracle and the
solution within 1e-8 of lexicographic GS; the mixed form), device and host
layouts byte-identical, a vector list in UG's mixed creation order, and the
entry points that take one vector type refusing two.
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


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _rhs(P, seed):
    lv = P.levels[-1]
    b = lv.b
    rhs = np.random.default_rng(seed).uniform(-1, 1, b * lv.n)
    rhs[np.repeat(lv.skip != 0, b)] = 0.0
    return rhs


@pytest.fixture(scope="module")
def p2():
    from ug_amd import problems
    P = problems.p2_uniform(6)          # 16,641 node + 49,408 edge vectors
    P.rhs = _rhs(P, 0x5EED0601)
    return P


def _ctx(P, smoother, damp, baselevel=2, setup=None, **kw):
    from ug_amd import ugmg
    ctx = ugmg.Context(0)
    if setup is not None:
        ctx.set_setup(setup)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=baselevel, smoother=smoother, damp=damp,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, **kw)
    ctx.preprocess()
    return ctx


def _kernels(P):
    from ug_amd import ugmg
    H = O.Hierarchy(P.levels)
    B = P.levels[-1].b
    cg, cj = _ctx(P, ugmg.SMOOTH_GS, 1.0), _ctx(P, ugmg.SMOOTH_JAC, 0.6)
    try:
        assert cg._ncomp() == 2 * B
        for l in range(3, P.top + 1):
            lv = P.levels[l]
            n2 = B * lv.n
            r = np.random.default_rng(l)
            x, y = r.uniform(-1, 1, n2), r.uniform(-1, 1, n2)
            xd = _dev(x)
            cg.dmatmul_minus(l, xd, _dev(y))
            assert _rel(_host(xd), O.dmatmul_minus(H, l, x, y)) < 1e-14, l
            np.testing.assert_allclose(cg.dnrm2x(l, _dev(x)), O.dnrm2x(H, l, x), rtol=1e-13)
            np.testing.assert_allclose(cj.dnrm2x(l, _dev(x)), O.dnrm2x(H, l, x), rtol=1e-13)
            np.testing.assert_allclose(cg.dnrm2x_noalloc(l, _dev(x)), O.dnrm2x(H, l, x), rtol=1e-13)
            np.testing.assert_allclose(cg.ddot(l, _dev(x), _dev(y)), O.ddot(H, l, x, y), rtol=1e-11)
            b = r.uniform(-1, 1, n2)
            xw, bw = O.smoother(H, l, O.JAC, 0.6, b)
            xd, bd = _dev(np.zeros(n2)), _dev(b)
            cj.jacobi_step(l, xd, bd)
            assert _rel(_host(xd), xw) < 1e-13 and _rel(_host(bd), bw) < 1e-12, l
            _, col = O.greedy_color(H, l)
            rank = (col.astype(np.int64) * lv.n + np.arange(lv.n)).astype(np.int32)
            xw, bw = O.smoother(H, l, O.GS, 1.0, b, rank=rank)
            xd, bd = _dev(np.zeros(n2)), _dev(b)
            cg.gs_step(l, xd, bd)
            assert _rel(_host(xd), xw) < 1e-12 and _rel(_host(bd), bw) < 1e-11, l
            nc2 = B * P.levels[l - 1].n
            bf = r.uniform(-1, 1, n2)
            cc = r.uniform(-1, 1, nc2)
            for ctx in (cg, cj):
                to_c = _dev(np.full(nc2, 5.0))
                ctx.restrict(l, to_c, _dev(bf))
                assert _rel(_host(to_c), O.restrict(H, l, bf, to_c=np.full(nc2, 5.0))) < 1e-14, l
                to_f = _dev(np.full(n2, 3.0))
                ctx.interpolate(l, to_f, _dev(cc), damp=0.7)
                assert _rel(_host(to_f), O.interpolate(H, l, cc, damp=0.7)) < 1e-15, l
        for smoother, ctx, kind, damp in ((O.MCGS, cg, "gs", 1.0), (O.JAC, cj, "jac", 0.6)):
            cw, bw = O.lmgc_iter(H, P.top, O.params(smoother=smoother, damp=damp, baselevel=2), P.rhs)
            c, bn = ctx.iter_host(P.top, P.rhs)
            assert _rel(c, cw) < 1e-11 and _rel(bn, bw) < 1e-10, kind
    finally:
        cg.close()
        cj.close()


def test_types_kernels(p2):
    _kernels(p2)


def test_types_kernels_in_creation_order():
    from ug_amd import problems
    P = problems.reordered(problems.p2_uniform(5), 0x5EED0602)
    assert not np.all(np.diff(P.levels[-1].vtype.astype(int)) >= 0)   # the types interleave
    P.rhs = _rhs(P, 0x5EED0603)
    _kernels(P)


def test_types_interpolation_reads_no_edge_skip_bits():
    from ug_amd import problems, ugmg
    P = problems.p2_uniform(5)
    f = P.levels[-1]
    e = np.flatnonzero((f.vtype == 1) & (f.skip == 0))[7]
    f.skip = f.skip.copy()
    f.skip[e] = 1
    P.rhs = _rhs(P, 0x5EED0604)
    H = O.Hierarchy(P.levels)
    c = P.levels[-2]
    cc = np.random.default_rng(1).uniform(1, 2, c.n)
    want = O.interpolate(H, P.top, cc)
    assert want[e] > 0
    ctx = _ctx(P, ugmg.SMOOTH_GS, 1.0)
    try:
        to_f = _dev(np.zeros(f.n))
        ctx.interpolate(P.top, to_f, _dev(cc))
        assert _rel(_host(to_f), want) < 1e-15
    finally:
        ctx.close()


def test_types_layouts_device_equals_host(p2):
    from ug_amd import ugmg
    cd = _ctx(p2, ugmg.SMOOTH_GS, 1.0, setup=ugmg.SETUP_DEVICE)
    ch = _ctx(p2, ugmg.SMOOTH_GS, 1.0, setup=ugmg.SETUP_HOST)
    try:
        for l in range(2, p2.top + 1):
            assert np.array_equal(cd.layout_digest(l), ch.layout_digest(l)), l
    finally:
        cd.close()
        ch.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_types_jacobi_solve_defects_per_type(p2, graph):
    from ug_amd import ugmg
    P = p2
    want = O.ls_solve(O.Hierarchy(P.levels), P.top, O.params(smoother=O.JAC, damp=0.6, baselevel=2), P.rhs,
                      red=1e-10, maxiter=200)
    ctx = _ctx(P, ugmg.SMOOTH_JAC, 0.6, use_graph=graph)
    try:
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=200)
    finally:
        ctx.close()
    assert res.converged and res.number_of_linear_iterations == want["nit"]
    assert d.shape == want["defects"].shape and d.shape[1] == 2
    assert np.max(np.abs(d - want["defects"]) / want["defects"]) < 1e-10
    assert _rel(x, want["x"]) < 1e-10


@pytest.mark.parametrize("precision", [0, 1])
def test_types_gs_solve(p2, precision):
    from ug_amd import ugmg
    P = p2
    H = O.Hierarchy(P.levels)
    lex = O.ls_solve(H, P.top, O.params(smoother=O.GS, damp=1.0, baselevel=2), P.rhs, red=1e-10, maxiter=200)
    mc = O.ls_solve(H, P.top, O.params(smoother=O.MCGS, damp=1.0, baselevel=2), P.rhs, red=1e-10, maxiter=200)
    ctx = _ctx(P, ugmg.SMOOTH_GS, 1.0, precision=precision, inner_cycles=2, use_graph=1)
    try:
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=200)
    finally:
        ctx.close()
    assert res.converged and d.shape[1] == 2
    assert np.all(np.array(res.last_defect[:2]) < 1e-10 * np.array(res.first_defect[:2]))
    assert np.max(np.abs(x - lex["x"])) < 1e-8 * np.max(np.abs(lex["x"]))
    if precision == 0:
        assert res.number_of_linear_iterations == mc["nit"]
        assert np.max(np.abs(d - mc["defects"]) / mc["defects"]) < 1e-9


def test_types_with_blocks_jacobi_solve():
    from ug_amd import problems, ugmg
    P = problems.p2_uniform(5, 2)       # 2 types x 2 components: 4 norm entries
    P.rhs = _rhs(P, 0x5EED0605)
    want = O.ls_solve(O.Hierarchy(P.levels), P.top, O.params(smoother=O.JAC, damp=0.6, baselevel=2), P.rhs,
                      red=1e-10, maxiter=200)
    ctx = _ctx(P, ugmg.SMOOTH_JAC, 0.6)
    try:
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=200)
    finally:
        ctx.close()
    assert res.converged and res.number_of_linear_iterations == want["nit"]
    assert d.shape == want["defects"].shape and d.shape[1] == 4
    assert np.max(np.abs(d - want["defects"]) / want["defects"]) < 1e-10


def test_types_refused_where_one_type_is_taken(p2):
    from ug_amd import problems, ugmg
    ctx = _ctx(p2, ugmg.SMOOTH_JAC, 0.6)
    try:
        with pytest.raises(ugmg.UgmgError) as e:
            ctx.solve_surface_host(p2.top, problems.surface_rhs(p2, 2, 1))
        assert e.value.code == -1
        with pytest.raises(ugmg.UgmgError) as e:
            ctx.bcgs_host(p2.top, p2.rhs)
        assert e.value.code == -1
    finally:
        ctx.close()


def test_types_through_the_numproc_layer(p2):
    from ug_amd import numproc as NP
    from ug_amd import ugmg
    P = p2
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    env = NP.NumProcEnv(ctx)
    x = env.vector("sol", np.zeros(P.levels[-1].n))
    env.vector("rhs", P.rhs.copy())
    env.create("transfer", "transfer"); env.init("transfer")
    env.create("jac", "jac"); env.init("jac", "$damp 0.6")
    env.create("ex", "ex"); env.init("ex")
    env.create("base", "ls"); env.init("base", "$red 1e-12 $m 1 $I ex $abslimit 1e-300")
    env.create("mgc", "lmgc"); env.init("mgc", "$S jac jac base $T transfer $n1 2 $n2 2 $b 2")
    env.create("mgs", "ls")
    assert env.init("mgs", "$A MAT $x sol $b rhs $m 200 $red 1e-10 $abslimit 1e-300 $I mgc $display full") \
        == NP.EXECUTABLE
    env.execute("mgs", "$i $d $r $s $p")
    r = env.result("mgs")
    out = env.output()
    env.close()
    ctx.close()
    want = O.ls_solve(O.Hierarchy(P.levels), P.top, O.params(smoother=O.JAC, damp=0.6, baselevel=2), P.rhs,
                      red=1e-10, maxiter=200)
    assert r.converged and r.number_of_linear_iterations == want["nit"]
    np.testing.assert_allclose(r.last_defect[:2], want["defects"][-1], rtol=1e-10)
    assert np.max(np.abs(x - want["x"])) < 1e-10 * np.max(np.abs(want["x"]))
    # DoPCR prints one line per norm entry: node ("u") and edge ("v") vectors
    import re
    assert len(re.findall(r"^      v: ", out, re.M)) == want["nit"] + 1


def test_types_typed_vec_scalars_through_the_numproc_layer(p2):
    """UG's typed VEC_SCALAR form (sc_read over ReadVecTypeDOUBLEs,
    np/udm/npscan.cc:783-855, 1211-1278): "$red n1e-6|k1e-10" gives node and
    edge norms their own reduction, "$damp n0.6|k0.6" one factor per type;
    a count per type that does not match the vectors is refused."""
    from ug_amd import numproc as NP
    from ug_amd import ugmg
    P = p2
    want = O.ls_solve(O.Hierarchy(P.levels), P.top, O.params(smoother=O.JAC, damp=0.6, baselevel=2), P.rhs,
                      red=[1e-6, 1e-10], maxiter=200)
    for red, ok in (("n1e-6|k1e-10", True), ("n1e-6", False)):
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(P)
        env = NP.NumProcEnv(ctx)
        x = env.vector("sol", np.zeros(P.levels[-1].n))
        env.vector("rhs", P.rhs.copy())
        env.create("transfer", "transfer"); env.init("transfer")
        env.create("jac", "jac"); env.init("jac", "$damp n0.6|k0.6")
        env.create("ex", "ex"); env.init("ex")
        env.create("base", "ls"); env.init("base", "$red 1e-12 $m 1 $I ex $abslimit 1e-300")
        env.create("mgc", "lmgc"); env.init("mgc", "$S jac jac base $T transfer $n1 2 $n2 2 $b 2")
        env.create("mgs", "ls")
        assert env.init("mgs", f"$A MAT $x sol $b rhs $m 200 $red {red} $abslimit 1e-300 $I mgc") \
            == NP.EXECUTABLE
        try:
            if ok:
                env.execute("mgs", "$i $d $r $s $p")
                r = env.result("mgs")
            else:
                with pytest.raises(ugmg.UgmgError):
                    env.execute("mgs", "$i $d $r $s $p")
        finally:
            env.close()
            ctx.close()
        if ok:
            assert r.converged and r.number_of_linear_iterations == want["nit"]
            np.testing.assert_allclose(r.last_defect[:2], want["defects"][-1], rtol=1e-10)
            assert np.max(np.abs(x - want["x"])) < 1e-10 * np.max(np.abs(want["x"]))


def test_types_with_blocks_damping_per_component():
    """$damp per component on node + edge vectors with 2 components: the
    factor of component j scales entry j of both types (dscalx with the
    same VEC_SCALAR value for (node, j) and (edge, j)); Jacobi and MC-GS
    solves against the oracle."""
    from ug_amd import problems, ugmg
    P = problems.p2_uniform(5, 2)
    P.rhs = _rhs(P, 0x5EED0606)
    H = O.Hierarchy(P.levels)
    for sm, osm, dm in ((ugmg.SMOOTH_JAC, O.JAC, [0.55, 0.65]), (ugmg.SMOOTH_GS, O.MCGS, [0.9, 1.0])):
        want = O.ls_solve(H, P.top, O.params(smoother=osm, damp=dm + dm, lmgc_damp=[0.95, 0.85] * 2,
                                             baselevel=2), P.rhs, red=1e-10, maxiter=200)
        ctx = _ctx(P, sm, dm, lmgc_damp=[0.95, 0.85])
        try:
            x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=200)
        finally:
            ctx.close()
        assert res.converged and res.number_of_linear_iterations == want["nit"], sm
        assert np.max(np.abs(d - want["defects"]) / want["defects"]) < 1e-9, sm
