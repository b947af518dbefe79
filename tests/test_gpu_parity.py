"""GPU parity: libugmg (through its C ABI) against the CPU oracle.

Sizes are those the oracle finishes in seconds: the 129x129 Q1 Poisson
hierarchy and the 17^3 / 33^3 jump-coefficient hierarchies.  Bars:
index arrays bit-exact; FP64 kernels within a few ulp-scaled relative error
of the oracle (the GPU contracts multiply-adds into FMAs, the oracle does
not); per-iteration defect norms of the Jacobi lmgc within 1e-10 relative;
multicolour GS final solution within 1e-8 of UG's lexicographic-GS solve.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a, dtype=torch.float64):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)
    torch.cuda.synchronize()   # the library runs on its own stream
    return t


def _full(n, val=0.0, dtype=torch.float64):
    t = torch.full((n,), val, dtype=dtype, device="cuda")
    torch.cuda.synchronize()
    return t


def _host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def _rel(a, b):
    d = np.max(np.abs(a - b))
    s = max(np.max(np.abs(b)), 1e-300)
    return d / s


def _ctx(problem, smoother, baselevel, damp, **kw):
    from ug_amd import ugmg
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(problem)
    ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=baselevel, smoother=smoother, damp=damp,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, **kw)
    ctx.preprocess()
    return ctx


def _rng(seed):
    return np.random.default_rng(seed)


@pytest.fixture(scope="module")
def ctx2d_jac(poisson2d):
    from ug_amd import ugmg
    c = _ctx(poisson2d, ugmg.SMOOTH_JAC, 0, 0.8)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx3d_gs(jump3d_17):
    from ug_amd import ugmg
    c = _ctx(jump3d_17, ugmg.SMOOTH_GS, 1, 1.0)
    yield c
    c.close()


def test_device_pattern_is_ug_pattern(poisson2d, ctx2d_jac, jump3d_17, ctx3d_gs):
    for P, ctx in ((poisson2d, ctx2d_jac), (jump3d_17, ctx3d_gs)):
        for l in range(ctx.params.baselevel, P.top + 1):
            rp, ci, va = ctx.pattern(l)
            lv = P.levels[l]
            assert np.array_equal(rp, lv.row_ptr)
            assert np.array_equal(ci, lv.col_idx)
            assert np.array_equal(va, lv.vals)


@pytest.mark.parametrize("which", ["2d", "3d"])
def test_dmatmul_minus_all_levels(which, poisson2d, ctx2d_jac, jump3d_17, ctx3d_gs):
    P, ctx = (poisson2d, ctx2d_jac) if which == "2d" else (jump3d_17, ctx3d_gs)
    H = O.Hierarchy(P.levels)
    for l in range(ctx.params.baselevel, P.top + 1):
        n = P.levels[l].n
        r = _rng(l)
        x, y = r.uniform(-1, 1, n), r.uniform(-1, 1, n)
        want = O.dmatmul_minus(H, l, x, y)
        xd, yd = _dev(x), _dev(y)
        ctx.dmatmul_minus(l, xd, yd)
        assert _rel(_host(xd), want) < 1e-14, l
        yd2 = _full(n)
        ctx.dmatmul(l, yd2, yd)
        assert _rel(_host(yd2), O.dmatmul(H, l, y)) < 1e-14, l
        # FP32 copy of A
        xs, ys = _dev(x, torch.float32), _dev(y, torch.float32)
        ctx.smatmul_minus(l, xs, ys)
        assert _rel(_host(xs), want) < 2e-5, l


def test_norms(poisson2d, ctx2d_jac):
    H = O.Hierarchy(poisson2d.levels)
    x = _rng(3).uniform(-1, 1, poisson2d.levels[-1].n)
    y = _rng(4).uniform(-1, 1, poisson2d.levels[-1].n)
    xd, yd = _dev(x), _dev(y)
    top = poisson2d.top
    assert abs(ctx2d_jac.dnrm2x(top, xd) - O.dnrm2x(H, top, x)[0]) < 1e-13 * np.linalg.norm(x)
    assert abs(ctx2d_jac.ddot(top, xd, yd) - float(x @ y)) < 1e-12 * np.linalg.norm(x) * np.linalg.norm(y)


def test_jacobi_step(poisson2d, ctx2d_jac):
    H = O.Hierarchy(poisson2d.levels)
    for l in range(1, poisson2d.top + 1):
        n = poisson2d.levels[l].n
        b = _rng(10 + l).uniform(-1, 1, n)
        xw, bw = O.smoother(H, l, O.JAC, 0.8, b)
        xd, bd = _full(n), _dev(b)
        ctx2d_jac.jacobi_step(l, xd, bd)
        assert _rel(_host(xd), xw) < 1e-15, l
        assert _rel(_host(bd), bw) < 1e-13, l


def test_colouring_matches_oracle(jump3d_17, ctx3d_gs):
    H = O.Hierarchy(jump3d_17.levels)
    lv = jump3d_17.levels
    for l in range(ctx3d_gs.params.baselevel + 1, jump3d_17.top + 1):
        col = ctx3d_gs.colors(l)
        nc, want = O.greedy_color(H, l)
        assert np.array_equal(col, want)
        assert nc == 8  # trilinear 27-point pattern: 2x2x2 parity classes
        rp, ci = lv[l].row_ptr, lv[l].col_idx
        rows = np.repeat(np.arange(lv[l].n), np.diff(rp))
        off = ci != rows
        assert not np.any(col[rows[off]] == col[ci[off]])


def test_gs_step_is_colour_ordered_lgs(jump3d_17, ctx3d_gs):
    H = O.Hierarchy(jump3d_17.levels)
    for l in range(ctx3d_gs.params.baselevel + 1, jump3d_17.top + 1):
        n = jump3d_17.levels[l].n
        _, col = O.greedy_color(H, l)
        rank = col.astype(np.int64) * n + np.arange(n)
        b = _rng(20 + l).uniform(-1, 1, n)
        xw, bw = O.smoother(H, l, O.GS, 1.0, b, rank=rank.astype(np.int32) if rank.max() < 2**31 else None)
        xd, bd = _full(n), _dev(b)
        ctx3d_gs.gs_step(l, xd, bd)
        assert _rel(_host(xd), xw) < 1e-13, l
        assert _rel(_host(bd), bw) < 1e-12, l


@pytest.mark.parametrize("which", ["2d", "3d"])
def test_transfer(which, poisson2d, ctx2d_jac, jump3d_17, ctx3d_gs):
    P, ctx = (poisson2d, ctx2d_jac) if which == "2d" else (jump3d_17, ctx3d_gs)
    H = O.Hierarchy(P.levels)
    for l in range(ctx.params.baselevel + 1, P.top + 1):
        nf, nc = P.levels[l].n, P.levels[l - 1].n
        bf = _rng(30 + l).uniform(-1, 1, nf)
        cc = _rng(40 + l).uniform(-1, 1, nc)
        want_r = O.restrict(H, l, bf)
        want_i = O.interpolate(H, l, cc, damp=0.7)
        to_c = _full(nc, 123.0)
        ctx.restrict(l, to_c, _dev(bf))
        assert _rel(_host(to_c), want_r) < 1e-14, l
        to_f = _full(nf, 7.0)
        ctx.interpolate(l, to_f, _dev(cc), damp=0.7)
        assert _rel(_host(to_f), want_i) < 1e-15, l


def test_base_solve_is_band_lu(jump3d_17, ctx3d_gs):
    H = O.Hierarchy(jump3d_17.levels)
    bl = ctx3d_gs.params.baselevel
    n = jump3d_17.levels[bl].n
    b = _rng(50).uniform(-1, 1, n)
    ex = O.Ex(H, bl, reorder=1)
    rc, xw, bw = ex.smooth(b)
    assert rc == 0
    cd, bd = _full(n), _dev(b)
    ctx3d_gs.base_solve(cd, bd)
    assert _rel(_host(cd), xw) < 1e-12
    assert np.max(np.abs(_host(bd) - bw)) < 1e-12 * np.max(np.abs(b))


@pytest.mark.parametrize("which", ["2d_jac", "3d_gs"])
def test_lmgc_iter(which, poisson2d, ctx2d_jac, jump3d_17, ctx3d_gs):
    if which == "2d_jac":
        P, ctx, p = poisson2d, ctx2d_jac, O.params(smoother=O.JAC, damp=0.8, baselevel=0)
    else:
        P, ctx, p = jump3d_17, ctx3d_gs, O.params(smoother=O.MCGS, damp=1.0, baselevel=1)
    H = O.Hierarchy(P.levels)
    cw, bw = O.lmgc_iter(H, P.top, p, P.rhs)
    cd = _full(P.levels[-1].n)
    bd = _dev(P.rhs)
    ctx.iter(P.top, cd, bd)
    assert _rel(_host(cd), cw) < 1e-12
    assert _rel(_host(bd), bw) < 1e-10


@pytest.mark.parametrize("graph", [0, 1])
def test_jacobi_solve_defects_match_ug(poisson2d, graph):
    """Per-iteration defect norms of ls+lmgc V(2,2) Jacobi 0.8 within 1e-10."""
    from ug_amd import ugmg
    ctx = _ctx(poisson2d, ugmg.SMOOTH_JAC, 0, 0.8, use_graph=graph)
    H = O.Hierarchy(poisson2d.levels)
    want = O.ls_solve(H, poisson2d.top, O.params(smoother=O.JAC, damp=0.8, baselevel=0), poisson2d.rhs,
                      red=1e-10, maxiter=50)
    n = poisson2d.levels[-1].n
    x, b = _full(n), _dev(poisson2d.rhs)
    res, defects = ctx.solve(poisson2d.top, x, b, red=1e-10, maxiter=50)
    d = np.array(defects)[:, 0]
    w = want["defects"][:, 0]
    assert res.converged and want["converged"]
    assert res.number_of_linear_iterations == want["nit"]
    assert len(d) == len(w)
    assert np.max(np.abs(d - w) / w) < 1e-10
    assert _rel(_host(x), want["x"]) < 1e-10
    ctx.close()


def test_jacobi_solve_3d_defects_match_ug(jump3d_33):
    from ug_amd import ugmg
    P = jump3d_33
    ctx = _ctx(P, ugmg.SMOOTH_JAC, 1, 0.8, use_graph=1)
    H = O.Hierarchy(P.levels)
    want = O.ls_solve(H, P.top, O.params(smoother=O.JAC, damp=0.8, baselevel=1), P.rhs, red=1e-10, maxiter=200)
    n = P.levels[-1].n
    x, b = _full(n), _dev(P.rhs)
    res, defects = ctx.solve(P.top, x, b, red=1e-10, maxiter=200)
    d = np.array(defects)[:, 0]
    w = want["defects"][:, 0]
    assert res.number_of_linear_iterations == want["nit"]
    assert np.max(np.abs(d - w) / w) < 1e-10
    ctx.close()


@pytest.mark.parametrize("bl,maxit", [(1, 1), (2, 2)])
def test_w_cycle_jacobi_defects_match_ug(jump3d_33, bl, maxit):
    """lmgc $g 2 (W-cycle): the base level is visited twice per visit of the
    level above, the second time on the defect the first leaves."""
    from ug_amd import ugmg
    P = jump3d_33
    H = O.Hierarchy(P.levels)
    want = O.ls_solve(H, P.top, O.params(smoother=O.JAC, damp=0.8, baselevel=bl, gamma=2, base_maxiter=maxit),
                      P.rhs, red=1e-10, maxiter=100)
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, gamma=2, baselevel=bl, smoother=ugmg.SMOOTH_JAC, damp=0.8,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=maxit, use_graph=1)
    ctx.preprocess()
    n = P.levels[-1].n
    x, b = _full(n), _dev(P.rhs)
    print(f"w-cycle bl={bl} maxit={maxit}", flush=True)
    res, defects = ctx.solve(P.top, x, b, red=1e-10, maxiter=100)
    d = np.array(defects)[:, 0]
    w = want["defects"][:, 0]
    assert res.number_of_linear_iterations == want["nit"]
    assert np.max(np.abs(d - w) / w) < 1e-10
    ctx.close()


def test_mcgs_solution_matches_ug_lexicographic_gs(jump3d_33):
    from ug_amd import ugmg
    P = jump3d_33
    H = O.Hierarchy(P.levels)
    want = O.ls_solve(H, P.top, O.params(smoother=O.GS, damp=1.0, baselevel=1), P.rhs, red=1e-10, maxiter=200)
    ctx = _ctx(P, ugmg.SMOOTH_GS, 1, 1.0, use_graph=1)
    n = P.levels[-1].n
    x, b = _full(n), _dev(P.rhs)
    res, defects = ctx.solve(P.top, x, b, red=1e-10, maxiter=200)
    assert res.converged
    xs = _host(x)
    assert np.linalg.norm(xs - want["x"]) / np.linalg.norm(want["x"]) < 1e-8
    # and the MC-GS iterates themselves equal the oracle's colour-ordered restatement
    mc = O.ls_solve(H, P.top, O.params(smoother=O.MCGS, damp=1.0, baselevel=1), P.rhs, red=1e-10, maxiter=200)
    d = np.array(defects)[:, 0]
    assert res.number_of_linear_iterations == mc["nit"]
    assert np.max(np.abs(d - mc["defects"][:, 0]) / mc["defects"][:, 0]) < 1e-9
    ctx.close()


@pytest.mark.parametrize("smoother", ["jac", "gs"])
def test_mixed_precision_reaches_fp64_accuracy(jump3d_33, smoother):
    """FP32 V-cycles inside FP64 refinement reach the same 1e-10 defect
    reduction; solution error against the FP64 oracle solve stated."""
    from ug_amd import ugmg
    P = jump3d_33
    sm = ugmg.SMOOTH_JAC if smoother == "jac" else ugmg.SMOOTH_GS
    damp = 0.8 if smoother == "jac" else 1.0
    ctx = _ctx(P, sm, 1, damp, precision=ugmg.MIXED, inner_cycles=2, use_graph=1)
    n = P.levels[-1].n
    x, b = _full(n), _dev(P.rhs)
    res, defects = ctx.solve(P.top, x, b, red=1e-10, maxiter=200)
    assert res.converged
    d = np.array(defects)[:, 0]
    assert d[-1] < 1e-10 * d[0]
    # true defect of the returned x, recomputed in FP64 on the host
    xs = _host(x)
    H = O.Hierarchy(P.levels)
    r = O.dmatmul_minus(H, P.top, P.rhs, xs)
    assert np.linalg.norm(r) < 2e-10 * np.linalg.norm(P.rhs)
    ref = O.ls_solve(H, P.top, O.params(smoother=O.GS, damp=1.0, baselevel=1), P.rhs, red=1e-13, maxiter=400)
    err = np.linalg.norm(xs - ref["x"]) / np.linalg.norm(ref["x"])
    print(f"mixed {smoother}: outer iterations {res.number_of_linear_iterations}, rel. error vs FP64 {err:.2e}")
    assert err < 1e-8
    ctx.close()


def test_mixed_coarse_inverse_precision(jump3d_33):
    """The FP32 cycle's base solve reads the FP64 dense inverse by default
    (the coarsest-grid solve kept in FP64); its FP32 copy (base_inv32) gives
    the same outer iterations and FP32 cycle count and the same final
    accuracy against a 1e-13 FP64 oracle solve."""
    from ug_amd import ugmg
    P = jump3d_33
    H = O.Hierarchy(P.levels)
    ref = O.ls_solve(H, P.top, O.params(smoother=O.GS, damp=1.0, baselevel=1), P.rhs, red=1e-13, maxiter=400)
    n = P.levels[-1].n
    out = {}
    for inv32 in (0, 1):
        ctx = _ctx(P, ugmg.SMOOTH_GS, 1, 1.0, precision=ugmg.MIXED, inner_cycles=7, use_graph=1,
                   base_inv32=inv32)
        x, b = _full(n), _dev(P.rhs)
        res, _ = ctx.solve(P.top, x, b, red=1e-10, maxiter=200)
        cycles, by = ctx.solve_stats()
        assert res.converged and by > 0
        xs = _host(x)
        r = O.dmatmul_minus(H, P.top, P.rhs, xs)
        out[inv32] = (res.number_of_linear_iterations, cycles, np.linalg.norm(r) / np.linalg.norm(P.rhs),
                      np.linalg.norm(xs - ref["x"]) / np.linalg.norm(ref["x"]))
        ctx.close()
    print(f"FP64 inverse {out[0]}, FP32 copy {out[1]} (outer its, FP32 cycles, true defect, error)")
    assert out[0][:2] == out[1][:2]
    for k in (0, 1):
        assert out[k][2] < 2e-10 and out[k][3] < 1e-8


def test_host_entry_points(poisson2d):
    from ug_amd import ugmg
    ctx = _ctx(poisson2d, ugmg.SMOOTH_JAC, 0, 0.8)
    H = O.Hierarchy(poisson2d.levels)
    c, b = ctx.iter_host(poisson2d.top, poisson2d.rhs)
    cw, bw = O.lmgc_iter(H, poisson2d.top, O.params(smoother=O.JAC, damp=0.8, baselevel=0), poisson2d.rhs)
    assert _rel(c, cw) < 1e-12 and _rel(b, bw) < 1e-10
    x, b, res, d = ctx.solve_host(poisson2d.top, poisson2d.rhs, red=1e-10)
    assert res.converged
    ctx.close()


def test_errors_are_reported(poisson2d):
    from ug_amd import ugmg
    ctx = ugmg.Context(0)
    with pytest.raises(ugmg.UgmgError) as e:
        ctx.preprocess(0)
    assert e.value.code in (-1, -3)
    ctx.set_hierarchy(poisson2d)
    ctx.set_params(baselevel=6)   # base level of 16641 vectors: refused
    with pytest.raises(ugmg.UgmgError):
        ctx.preprocess()
    ctx.close()


@pytest.mark.parametrize("bl,maxit", [(3, 1), (4, 1), (2, 2)])
def test_base_levels_and_base_iterations(jump3d_33, bl, maxit):
    """Dense-inverse base solve at several base levels, and a base ls with
    $m 2 (two ex applications), against the oracle's band-LU restatement."""
    from ug_amd import ugmg
    P = jump3d_33
    H = O.Hierarchy(P.levels)
    want = O.ls_solve(H, P.top, O.params(smoother=O.MCGS, damp=1.0, baselevel=bl, base_maxiter=maxit,
                                         base_red=1e-3), P.rhs, red=1e-10, maxiter=200)
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=bl, smoother=ugmg.SMOOTH_GS, damp=1.0,
                   base_abslimit=1e-300, base_red=1e-3, base_maxiter=maxit, use_graph=1)
    ctx.preprocess()
    n = P.levels[-1].n
    x, b = _full(n), _dev(P.rhs)
    res, defects = ctx.solve(P.top, x, b, red=1e-10, maxiter=200)
    d = np.array(defects)[:, 0]
    assert res.number_of_linear_iterations == want["nit"]
    assert np.max(np.abs(d - want["defects"][:, 0]) / want["defects"][:, 0]) < 1e-9
    ctx.close()


def test_fused_coarse_cycle_matches(poisson2d, monkeypatch):
    """The single-workgroup coarse-cycle kernel (levels up to UGMG_FUSE_ROWS
    vectors) gives the oracle's defects, W-cycle (gamma 2) included."""
    from ug_amd import ugmg
    H = O.Hierarchy(poisson2d.levels)
    monkeypatch.setenv("UGMG_FUSE_ROWS", "5000")
    for gamma, sm in ((1, ugmg.SMOOTH_JAC), (2, ugmg.SMOOTH_GS)):
        osm = O.JAC if sm == ugmg.SMOOTH_JAC else O.MCGS
        damp = 0.8 if sm == ugmg.SMOOTH_JAC else 1.0
        want = O.ls_solve(H, poisson2d.top, O.params(smoother=osm, damp=damp, baselevel=0, gamma=gamma),
                          poisson2d.rhs, red=1e-10, maxiter=50)
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(poisson2d)
        ctx.set_params(nu1=2, nu2=2, gamma=gamma, baselevel=0, smoother=sm, damp=damp, base_abslimit=1e-300,
                       use_graph=1)
        ctx.preprocess()
        n = poisson2d.levels[-1].n
        x, b = _full(n), _dev(poisson2d.rhs)
        res, defects = ctx.solve(poisson2d.top, x, b, red=1e-10, maxiter=50)
        d = np.array(defects)[:, 0]
        assert res.number_of_linear_iterations == want["nit"], (gamma, sm)
        assert np.max(np.abs(d - want["defects"][:, 0]) / want["defects"][:, 0]) < 1e-10
        ctx.close()


@pytest.mark.parametrize("precision,nu", [(0, 2), (1, 3), (0, 1)])
def test_folded_gs_defects_are_bitwise_the_separate_passes(jump3d_17, monkeypatch, precision, nu):
    """Folding each GS step's defect into the next step's colour sweeps
    (the default), and the post-smoother's colour-0 sweep into the SpMV
    that applies the coarse-grid correction, change no bit of the cycle
    against separate passes (UGMG_GS_FUSE=0), in FP64 and in the FP32
    cycle; the folded sweep gives the same bits with its two row products on
    two wavefronts per slice (the default) or on one (UGMG_GS_SPLIT=0); and
    the list
    order (UGMG_GS_WINDOW=1) gives the device order's results to rounding."""
    from ug_amd import ugmg
    P = jump3d_17
    n = P.levels[-1].n
    b0 = P.rhs.copy()
    out = {}
    legs = (("fused", {}), ("split", {"UGMG_GS_FUSE": "0"}), ("onewave", {"UGMG_GS_SPLIT": "0"}),
            ("list", {"UGMG_GS_WINDOW": "1"}))
    for tag, env in legs:
        for k in ("UGMG_GS_FUSE", "UGMG_GS_WINDOW", "UGMG_GS_SPLIT"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(P)
        ctx.set_params(nu1=nu, nu2=nu, gamma=1, baselevel=1, smoother=ugmg.SMOOTH_GS, damp=1.0,
                       base_abslimit=1e-300, precision=precision, inner_cycles=1)
        ctx.preprocess()
        out[tag] = ctx.iter_host(P.top, b0)
        ctx.close()
    assert np.array_equal(out["fused"][0], out["split"][0]) and np.array_equal(out["fused"][1], out["split"][1])
    assert np.array_equal(out["fused"][0], out["onewave"][0]) and np.array_equal(out["fused"][1], out["onewave"][1])
    tol = 1e-12 if precision == 0 else 1e-5
    assert _rel(out["fused"][0], out["list"][0]) < tol


def test_server_results_equal_standalone_solves():
    """Continuous batching: four requests of different sizes and kinds
    admitted together (two slots, so two join while others run) give exactly
    the solutions of stand-alone solves."""
    from ug_amd import problems, ugmg
    probs = [problems.q1(2, 2, 5, problems.KAPPA_ONE, seed=1), problems.jump3d(4, seed=2),
             problems.q1(2, 2, 6, problems.KAPPA_ONE, seed=3), problems.jump3d(5, seed=4)]
    kw = dict(nu1=2, nu2=2, gamma=1, baselevel=1, smoother=ugmg.SMOOTH_GS, damp=1.0, base_abslimit=1e-300,
              use_graph=1)
    alone = []
    for P in probs:
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(P)
        ctx.set_params(**kw)
        ctx.preprocess()
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-8, maxiter=200)
        assert res.converged
        alone.append((x, res.number_of_linear_iterations))
        ctx.close()
    srv = ugmg.Server(0, budget_bytes=8e9, max_slots=2)
    ids = [srv.submit(P, ugmg.default_params(**kw), red=1e-8, maxiter=200) for P in probs]
    srv.drain(len(ids))
    for P, rid, (xa, nit) in zip(probs, ids, alone):
        st, res, times = srv.status(rid)
        assert st == ugmg.REQ_DONE and res.converged
        assert res.number_of_linear_iterations == nit
        assert times[0] <= times[1] <= times[2]
        xs = srv.result(rid, P.levels[-1].n)
        assert np.array_equal(xs, xa)
    srv.close()


def test_server_batches_same_shape_solves_bit_identically():
    """Six requests of two shapes (the 17^3 jump-coefficient hierarchy with
    three kappa draws, so three matrices of one pattern, and 65^2 Poisson
    with three right-hand sides), four slots: each shape runs as one
    resident group whose iterations are single batched launch sequences
    over its occupied slots, the last two requests join groups that are
    already iterating, and every result equals its stand-alone solve bit
    for bit."""
    from ug_amd import problems, ugmg
    probs = [problems.jump3d(4, seed=11), problems.q1(2, 2, 5, problems.KAPPA_ONE, seed=21),
             problems.jump3d(4, seed=12), problems.q1(2, 2, 5, problems.KAPPA_ONE, seed=22),
             problems.jump3d(4, seed=13), problems.q1(2, 2, 5, problems.KAPPA_ONE, seed=23)]
    kw = dict(nu1=2, nu2=2, gamma=1, baselevel=1, smoother=ugmg.SMOOTH_GS, damp=1.0, base_abslimit=1e-300,
              use_graph=1)
    alone = []
    for P in probs:
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(P)
        ctx.set_params(**kw)
        ctx.preprocess()
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-8, maxiter=200)
        assert res.converged
        alone.append((x, res.number_of_linear_iterations))
        ctx.close()
    assert not np.array_equal(probs[0].levels[-1].vals, probs[2].levels[-1].vals)
    srv = ugmg.Server(0, budget_bytes=8e9, max_slots=4)
    ids = [srv.submit(P, ugmg.default_params(**kw), red=1e-8, maxiter=200) for P in probs]
    srv.drain(len(ids))
    groups, launches, columns = srv.stats()
    print(f"groups {groups}, group iterations {launches}, solve iterations {columns}")
    assert groups == 2
    assert columns == sum(nit for _, nit in alone)
    assert columns > launches          # launches carried several solves
    for P, rid, (xa, nit) in zip(probs, ids, alone):
        st, res, times = srv.status(rid)
        assert st == ugmg.REQ_DONE and res.converged
        assert res.number_of_linear_iterations == nit
        assert np.array_equal(srv.result(rid, P.levels[-1].n), xa)
    srv.close()


def test_server_single_slot_groups_reuse_the_layout():
    """max_slots = 1: each group holds one member, which later requests of
    the same shape replace through the group's index maps (the group's
    layout is built once); results equal stand-alone solves bit for bit."""
    from ug_amd import problems, ugmg
    probs = [problems.jump3d(4, seed=31), problems.jump3d(4, seed=32), problems.jump3d(4, seed=33)]
    kw = dict(nu1=2, nu2=2, gamma=1, baselevel=1, smoother=ugmg.SMOOTH_GS, damp=1.0, base_abslimit=1e-300,
              use_graph=1)
    alone = []
    for P in probs:
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(P)
        ctx.set_params(**kw)
        ctx.preprocess()
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-8, maxiter=200)
        assert res.converged
        alone.append((x, res.number_of_linear_iterations))
        ctx.close()
    srv = ugmg.Server(0, budget_bytes=8e9, max_slots=1)
    ids = [srv.submit(P, ugmg.default_params(**kw), red=1e-8, maxiter=200) for P in probs]
    srv.drain(len(ids))
    groups, launches, columns = srv.stats()
    assert groups == 1 and columns == launches == sum(nit for _, nit in alone)
    for P, rid, (xa, nit) in zip(probs, ids, alone):
        st, res, _ = srv.status(rid)
        assert st == ugmg.REQ_DONE and res.converged and res.number_of_linear_iterations == nit
        assert np.array_equal(srv.result(rid, P.levels[-1].n), xa)
    srv.close()


def test_server_budget_rejects_oversized_request():
    from ug_amd import problems, ugmg
    P = problems.jump3d(4, seed=5)
    srv = ugmg.Server(0, budget_bytes=1e5, max_slots=2)
    rid = srv.submit(P, ugmg.default_params(baselevel=1, base_abslimit=1e-300), red=1e-8)
    srv.drain(1)
    st, res, _ = srv.status(rid)
    assert st == ugmg.REQ_FAILED and res.error_code == -6
    srv.close()
