"""A UG script through the numproc layer on the GPU: the configuration of
UG's config-1 run (ls + lmgc V(2,2), jac $damp 0.8, ex base) executed with
npexecute's flags gives the oracle's solve; bcgs with an lmgc $I likewise."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_ls_script_is_the_oracle_solve(poisson2d):
    from ug_amd import numproc as NP
    from ug_amd import ugmg
    P = poisson2d
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    env = NP.NumProcEnv(ctx)
    x = env.vector("sol", np.zeros(P.levels[-1].n))
    b = env.vector("rhs", P.rhs)
    env.create("transfer", "transfer"); env.init("transfer")
    env.create("jac", "jac"); env.init("jac", "$damp 0.8")
    env.create("ex", "ex"); env.init("ex")
    env.create("base", "ls"); env.init("base", "$red 1e-12 $m 1 $I ex $abslimit 1e-300")
    env.create("mgc", "lmgc"); env.init("mgc", "$S jac jac base $T transfer $n1 2 $n2 2 $b 0 $graph 1")
    env.create("mgs", "ls")
    assert env.init("mgs", "$A MAT $x sol $b rhs $m 50 $red 1e-10 $abslimit 1e-300 $I mgc") == NP.EXECUTABLE
    print("npexecute mgs", flush=True)
    env.execute("mgs", "$i $d $r $s $p")
    r = env.result("mgs")
    want = O.ls_solve(O.Hierarchy(P.levels), P.top, O.params(smoother=O.JAC, damp=0.8, baselevel=0), P.rhs,
                      red=1e-10, maxiter=50)
    assert r.converged and r.number_of_linear_iterations == want["nit"]
    assert abs(r.last_defect[0] - want["defects"][-1, 0]) < 1e-10 * want["defects"][-1, 0]
    assert np.max(np.abs(x - want["x"])) < 1e-10 * np.max(np.abs(want["x"]))
    assert np.max(np.abs(b - want["b"])) < 1e-9 * np.max(np.abs(P.rhs))
    env.close()
    ctx.close()


def test_bcgs_script_with_two_rhs():
    from ug_amd import numproc as NP
    from ug_amd import problems, ugmg
    P = problems.p1_cd(5, 1e-2, 0.7, seed=3)
    n = P.levels[-1].n
    R = np.stack([P.rhs, np.where(P.levels[-1].skip == 1, 0.0, np.random.default_rng(2).uniform(-1, 1, n))])
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    env = NP.NumProcEnv(ctx)
    X = env.vector("x", np.zeros((2, n)))
    B = env.vector("b", R.copy())
    env.create("transfer", "transfer"); env.init("transfer")
    env.create("gs", "gs"); env.init("gs")
    env.create("ex", "ex"); env.init("ex")
    env.create("base", "ls"); env.init("base", "$red 1e-12 $m 1 $I ex $abslimit 1e-300")
    env.create("mgc", "lmgc"); env.init("mgc", "$S gs gs base $T transfer $n1 1 $n2 1 $b 0")
    env.create("k", "bcgs")
    assert env.init("k", "$A MAT $x x $b b $m 100 $red 1e-10 $abslimit 1e-300 $I mgc $nrhs 2") == NP.EXECUTABLE
    env.execute("k", "$i $d $r $s")
    H = O.Hierarchy(P.levels)
    for j in range(2):
        want = O.bcgs_solve(H, P.top, O.params(smoother=O.MCGS, damp=1.0, nu1=1, nu2=1), R[j], red=1e-10,
                            maxiter=100)
        r = env.result("k", j)
        assert r.converged and r.number_of_linear_iterations == want["nit"]
        assert np.max(np.abs(X[j] - want["x"])) < 1e-8 * np.max(np.abs(want["x"]))
    env.close()
    ctx.close()


def _ug_display(name, defects, level, nit):
    """LinearSolver's display text (np/procs/ls.cc:666-744 with DoPCR,
    np/udm/pcr.cc:505-606, full display, one component), times elided."""
    head = ("*" * ((50 - len(name)) // 2 - 1) + " " + name + " ").ljust(50, "*")
    lines = ["", head]
    lines.append(" %-3d  %c: %-12.7e   %-12.7s" % (0, "u", defects[0], "---"))
    for k in range(1, len(defects)):
        lines.append(" %-3d  %c: %-12.7e   %-12.7e" % (k, "u", defects[k], defects[k] / defects[k - 1]))
    lines.append("")
    nb = len(defects)
    lines.append(" %-3d avg:  %c: %-12.7e   %-12.7e   %-12.7e"
                 % (nb - 1, "u", defects[0], defects[-1], (defects[-1] / defects[0]) ** (1.0 / (nb - 1))))
    lines.append("")
    return "\n".join(lines) + "\n", "LS  : L=%2d N=%2d TSOLVE=" % (level, nit)


def test_ls_display_full_is_ugs_text(poisson2d):
    from ug_amd import numproc as NP
    from ug_amd import ugmg
    P = poisson2d
    # the defect sequence of the same solve, through the C ABI's callback
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, baselevel=0, smoother=ugmg.SMOOTH_JAC, damp=0.8, base_abslimit=1e-300,
                   base_red=1e-12, base_maxiter=1)
    ctx.preprocess()
    _, _, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=50)
    defects = [row[0] for row in d]
    env = NP.NumProcEnv(ctx)
    env.vector("sol", np.zeros(P.levels[-1].n))
    env.vector("rhs", P.rhs)
    env.create("transfer", "transfer"); env.init("transfer")
    env.create("jac", "jac"); env.init("jac", "$damp 0.8")
    env.create("ex", "ex"); env.init("ex")
    env.create("base", "ls"); env.init("base", "$red 1e-12 $m 1 $I ex $abslimit 1e-300")
    env.create("mgc", "lmgc"); env.init("mgc", "$S jac jac base $T transfer $n1 2 $n2 2 $b 0")
    env.create("mgs", "ls")
    env.init("mgs", "$A MAT $x sol $b rhs $m 50 $red 1e-10 $abslimit 1e-300 $I mgc $display full")
    env.execute("mgs", "$i $d $r $s")
    out = env.output()
    body, last = _ug_display("mgs", defects, P.top, res.number_of_linear_iterations)
    assert out.startswith(body), out[:400]
    tail = out[len(body):]
    assert tail.startswith(last) and tail.endswith("\n") and "TIT=" in tail
    assert env.output() == ""                        # handed out once
    env.init("mgs", "$A MAT $x sol $b rhs $m 50 $red 1e-10 $abslimit 1e-300 $I mgc $display red")
    env.vector("sol", np.zeros(P.levels[-1].n))
    env.execute("mgs", "$d $r $s")
    red = env.output()                               # head, averages and the LS line only
    assert " 1    u:" not in red and " avg:  u:" in red and "LS  : L=" in red
    env.close()
    ctx.close()


def test_ls_script_on_a_locally_refined_hierarchy():
    """The same ls script on a red-green refined hierarchy: npexecute runs
    LinearSolver ON_SURFACE over the registered levels and gives the
    oracle's surface solve; bcgs refuses such a hierarchy."""
    from ug_amd import numproc as NP
    from ug_amd import problems, ugmg
    Q = problems.p1_local(5, 2)
    bl = 1
    fs = problems.surface_rhs(Q, bl, 4)
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(Q)
    env = NP.NumProcEnv(ctx)
    x = env.vector("sol", np.zeros(Q.levels[-1].n))
    b = env.vector("rhs", fs[-1])
    xs = {l: env.level_vector("sol", l, np.zeros(Q.levels[l].n)) for l in range(bl, Q.top)}
    bs = {l: env.level_vector("rhs", l, fs[l]) for l in range(bl, Q.top)}
    env.create("transfer", "transfer"); env.init("transfer")
    env.create("jac", "jac"); env.init("jac", "$damp 0.8")
    env.create("ex", "ex"); env.init("ex")
    env.create("base", "ls"); env.init("base", "$red 1e-12 $m 1 $I ex $abslimit 1e-300")
    env.create("mgc", "lmgc"); env.init("mgc", f"$S jac jac base $T transfer $n1 2 $n2 2 $b {bl}")
    env.create("mgs", "ls")
    assert env.init("mgs", "$A MAT $x sol $b rhs $m 50 $red 1e-10 $abslimit 1e-300 $I mgc $display red") \
        == NP.EXECUTABLE
    print("npexecute mgs (locally refined)", flush=True)
    env.execute("mgs", "$i $d $r $s")
    r = env.result("mgs")
    text = env.output()
    want = O.ls_solve_surface(O.Hierarchy(Q.levels), Q.top, O.params(smoother=O.JAC, damp=0.8, baselevel=bl),
                              fs, red=1e-10, maxiter=50)
    assert r.converged and r.number_of_linear_iterations == want["nit"]
    assert abs(r.last_defect[0] - want["defects"][-1, 0]) < 1e-10 * want["defects"][-1, 0]
    assert np.max(np.abs(x - want["x"][-1])) < 1e-10 * np.max(np.abs(want["x"][-1]))
    for l in range(bl, Q.top):
        assert np.max(np.abs(xs[l] - want["x"][l])) < 1e-10 * np.max(np.abs(want["x"][l]))
        # b is the defect on the surface rows; the covered rows are the cycle's scratch
        m = problems.fine_grid_dof(Q.levels[l])
        assert np.max(np.abs(bs[l][m] - want["b"][l][m]), initial=0.0) < 1e-9
    assert "LS  : L=" in text
    env.create("k", "bcgs")
    assert env.init("k", "$A MAT $x sol $b rhs $m 10 $red 1e-8 $I mgc") == NP.EXECUTABLE
    with pytest.raises(Exception):
        env.execute("k", "$s")
    env.execute("mgs", "$p")
    env.close()
    ctx.close()
