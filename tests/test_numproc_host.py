"""The numproc layer on the CPU (no GPU touched): UG's option syntax,
defaults and returned states of npinit for the classes of the solve path
(LinearSolverInit ls.cc:771-797, NPLinearSolverInit ls.cc:382-407,
BCGSInit ls.cc:1750-1800, LmgcInit iter.cc:7613-7656, SmootherInit
iter.cc:763-790, EXInit iter.cc:8248-8260), and the parameters the chain
hands to libugmg."""
import numpy as np
import pytest

from ug_amd import numproc as NP
from ug_amd import ugmg


def script(env, smoother="gs", damp="1.0", extra_mg=""):
    env.create("transfer", "transfer")
    assert env.init("transfer") == NP.EXECUTABLE
    for s in ("pre", "post"):
        env.create(s, smoother)
        assert env.init(s, f"$damp {damp}") == NP.ACTIVE          # no $A $c $r
    env.create("ex", "ex")
    assert env.init("ex") == NP.ACTIVE
    env.create("base", "ls")
    assert env.init("base", "$red 1e-12 $m 1 $I ex $abslimit 1e-300") == NP.ACTIVE   # no $x $b $A
    env.create("mgc", "lmgc")
    return env.init("mgc", f"$S pre post base $T transfer $n1 2 $n2 2 $b 4 $damp 1.0 {extra_mg}")


def test_script_chain_and_params():
    env = NP.NumProcEnv(None)
    assert script(env) == NP.ACTIVE
    env.vector("sol", np.zeros(10))
    env.vector("rhs", np.ones(10))
    env.create("mgs", "ls")
    assert env.init("mgs", "$A MAT $x sol $b rhs $m 100 $red 1e-10 $I mgc") == NP.EXECUTABLE
    p = env.params("mgs")
    assert (p.nu1, p.nu2, p.gamma, p.baselevel, p.smoother) == (2, 2, 1, 4, ugmg.SMOOTH_GS)
    assert p.damp[0] == 1.0 and p.lmgc_damp[0] == 1.0 and p.base_damp[0] == 1.0
    assert p.base_maxiter == 1 and p.base_red[0] == 1e-12 and p.base_abslimit[0] == 1e-300
    assert p.ex_reorder == 1 and p.precision == ugmg.FP64 and p.inner_cycles == 1
    # execution needs a context
    with pytest.raises(ugmg.UgmgError):
        env.execute("mgs", "$i $s")


def test_defaults_and_states():
    env = NP.NumProcEnv(None)
    env.create("t", "transfer")
    env.init("t")
    env.create("j", "jac")
    assert env.init("j") == NP.ACTIVE
    env.create("ex", "ex")
    env.init("ex", "$o 0")
    env.create("base", "ls")
    env.init("base", "$m 2 $I ex $red 1e-6")
    env.create("mg", "lmgc")
    assert env.init("mg", "$S j j base $T t") == NP.ACTIVE
    env.create("s", "ls")
    env.vector("x", np.zeros(4))
    env.vector("b", np.zeros(4))
    assert env.init("s", "$x x $b b $A A $m 5 $I mg") == NP.ACTIVE      # $red missing
    assert env.init("s", "$x x $b b $A A $m 5 $I mg $red 1e-8") == NP.EXECUTABLE
    p = env.params("s")
    # LmgcInit defaults g = n1 = n2 = 1, b = 0; lmgc and smoother $damp 1
    assert (p.gamma, p.nu1, p.nu2, p.baselevel) == (1, 1, 1, 0)
    assert p.damp[2] == 1.0 and p.lmgc_damp[1] == 1.0
    assert p.ex_reorder == 0 and p.base_maxiter == 2 and p.base_red[0] == 1e-6
    assert p.base_abslimit[0] == 1e-10            # ABS_LIMIT (ls.cc:66)
    assert p.smoother == ugmg.SMOOTH_JAC


def test_error_behaviour():
    env = NP.NumProcEnv(None)
    assert script(env) == NP.ACTIVE
    env.create("a", "ls")
    assert env.init("a", "$red 1e-8 $I mgc") == NP.NOT_ACTIVE        # $m missing
    assert env.init("a", "$m 3 $red 1e-8 $I nothere") == NP.NOT_ACTIVE
    env.create("k", "bcgs")
    assert env.init("k", "$m 3 $R -1 $red 1e-8") == NP.NOT_ACTIVE    # $R < 0
    assert env.init("k", "$m 3 $red 1e-8 $weight 2") == NP.NOT_ACTIVE
    assert env.init("k", "$m 3 $red 1e-8") == NP.ACTIVE               # no vectors; no $I is allowed
    env.create("m2", "lmgc")
    assert env.init("m2", "$S pre post base") == NP.NOT_ACTIVE        # $T missing
    assert env.init("m2", "$S pre post base $T transfer $precision half") == NP.NOT_ACTIVE
    # ReadArgvINT matches the whole option name: "$nn1 5" is not $n1
    assert env.init("m2", "$S pre post base $T transfer $nn1 5") == NP.ACTIVE
    with pytest.raises(ugmg.UgmgError):
        env.create("m2", "lmgc")                                      # name taken
    with pytest.raises(ugmg.UgmgError):
        env.create("q", "cg")                                         # not a class of this path


def test_pre_and_post_smoother_objects():
    """$S pre post base (LmgcInit iter.cc:7613-7656): the post object's class
    and $damp are kept when they differ from the pre object's; an ex or
    lmgc object as a smoother is refused."""
    env = NP.NumProcEnv(None)
    script(env)                               # pre and post: gs, $damp 1.0
    env.vector("x", np.zeros(2))
    env.vector("b", np.zeros(2))
    for name, cls, damp, want in (("p1", "gs", "1.0", (0, None)), ("p2", "gs", "0.7", (1, ugmg.SMOOTH_GS)),
                                  ("p3", "jac", "1.0", (1, ugmg.SMOOTH_JAC))):
        env.create(name, cls)
        env.init(name, f"$damp {damp}")
        env.create("mg" + name, "lmgc")
        env.init("mg" + name, f"$S pre {name} base $T transfer")
        env.create("s" + name, "ls")
        env.init("s" + name, f"$x x $b b $A A $m 5 $I mg{name} $red 1e-8")
        p = env.params("s" + name)
        assert p.smoother == ugmg.SMOOTH_GS and p.damp[0] == 1.0
        assert p.post_set == want[0]
        if want[0]:
            assert p.post_smoother == want[1] and p.post_damp[0] == float(damp)
    env.create("mgx", "lmgc")
    env.init("mgx", "$S pre ex base $T transfer")
    env.create("sx", "ls")
    env.init("sx", "$x x $b b $A A $m 5 $I mgx $red 1e-8")
    with pytest.raises(ugmg.UgmgError):
        env.params("sx")


def test_gpu_options():
    env = NP.NumProcEnv(None)
    assert script(env, extra_mg="$precision mixed $inner 3 $graph 0") == NP.ACTIVE
    env.vector("x", np.zeros(2))
    env.vector("b", np.zeros(2))
    env.create("s", "bcgs")
    assert env.init("s", "$x x $b b $A A $m 50 $R 10 $I mgc $red 1e-8 $nrhs 4") == NP.EXECUTABLE
    p = env.params("s")
    assert p.precision == ugmg.MIXED and p.inner_cycles == 3 and p.use_graph == 0


def test_ex_float_option_reaches_the_parameters():
    """EXInit's $f (np/procs/iter.cc:8253): FLOAT band LU of the base level."""
    from ug_amd import numproc as NP
    for opt, want in (("$o 1", 0), ("$o 1 $f", 1)):
        env = NP.NumProcEnv(None)
        x = np.zeros(9)
        env.vector("sol", x)
        env.vector("rhs", x.copy())
        for name, cls, o in [("base", "ex", opt), ("bs", "ls", "$m 1 $red 1e-12 $I base"),
                             ("sm", "jac", "$damp 0.8"), ("tr", "transfer", ""),
                             ("mgc", "lmgc", "$S sm sm bs $T tr"),
                             ("mgs", "ls", "$x sol $b rhs $m 10 $red 1e-10 $I mgc")]:
            env.create(name, cls)
            env.init(name, o)
        assert env.params("mgs").ex_float == want
        env.close()


def test_typed_damp_per_component():
    """sc_read over ReadVecTypeDOUBLEs (np/udm/npscan.cc:783-855, 1211-1278):
    "$damp n0.5:0.6 0.7" gives the node components their own factors; one
    number gives every entry the same; a token without a type letter next
    to typed ones, or a two-letter type, is refused."""
    env = NP.NumProcEnv(None)
    assert script(env, damp="n0.5:0.6 0.7", extra_mg="") == NP.ACTIVE
    env.create("ex2", "ex")
    assert env.init("ex2", "$damp n0.9:0.8:0.7") == NP.ACTIVE
    env.create("base2", "ls")
    env.init("base2", "$red 1e-12 $m 1 $I ex2")
    env.create("mg2", "lmgc")
    env.init("mg2", "$S pre post base2 $T transfer $damp n1.0:0.5:0.25")
    env.vector("x", np.zeros(3))
    env.vector("b", np.zeros(3))
    env.create("s", "ls")
    assert env.init("s", "$x x $b b $A A $m 5 $I mg2 $red 1e-8") == NP.EXECUTABLE
    p = env.params("s")
    assert list(p.damp)[:3] == [0.5, 0.6, 0.7]
    assert list(p.base_damp)[:3] == [0.9, 0.8, 0.7]
    assert list(p.lmgc_damp)[:3] == [1.0, 0.5, 0.25]
    # node and edge values that agree are one factor per component
    env.create("g", "gs")
    assert env.init("g", "$damp n0.5:0.6|k0.5:0.6") == NP.ACTIVE
    env.create("mg3", "lmgc")
    env.init("mg3", "$S g g base2 $T transfer")
    env.create("s3", "ls")
    env.init("s3", "$x x $b b $A A $m 5 $I mg3 $red 1e-8")
    with pytest.raises(ugmg.UgmgError):      # ex2 gives 3 values, g 2: other counts per type
        env.params("s3")
    env.create("ex4", "ex")
    env.init("ex4", "$damp n0.9:0.8")
    env.create("base4", "ls")
    env.init("base4", "$red 1e-12 $m 1 $I ex4")
    env.create("mg4", "lmgc")
    env.init("mg4", "$S g g base4 $T transfer")
    env.create("s4", "ls")
    env.init("s4", "$x x $b b $A A $m 5 $I mg4 $red 1e-8")
    p = env.params("s4")
    assert list(p.damp)[:2] == [0.5, 0.6] and list(p.base_damp)[:2] == [0.9, 0.8]
    # node and edge factors that differ are refused (one per component)
    env.create("g5", "gs")
    env.init("g5", "$damp n0.5|k0.6")
    env.create("mg5", "lmgc")
    env.init("mg5", "$S g5 g5 base2 $T transfer")
    env.create("s5", "ls")
    env.init("s5", "$x x $b b $A A $m 5 $I mg5 $red 1e-8")
    with pytest.raises(ugmg.UgmgError):
        env.params("s5")
    # malformed forms
    env.create("bad", "jac")
    assert env.init("bad", "$damp 0.5|n0.6") == NP.NOT_ACTIVE       # type letter missing
    assert env.init("bad", "$damp nk0.5") == NP.NOT_ACTIVE          # two-letter type
    assert env.init("bad", "$damp 0.5 0.6") == NP.NOT_ACTIVE        # several values, no type
    assert env.init("bad", "$damp n0.5:x") == NP.NOT_ACTIVE         # unreadable value
    assert env.init("bad", "$damp 0.75") == NP.ACTIVE


def test_base_ls_over_jac():
    """The base ls may iterate with a jac object (LinearSolver over Smoother
    on the base level): base_iter BASE_JAC, the jac object's $damp as the
    base damping; a gs object there is refused."""
    env = NP.NumProcEnv(None)
    script(env)
    env.create("bj", "jac")
    env.init("bj", "$damp 0.65")
    env.create("bbase", "ls")
    env.init("bbase", "$red 1e-3 $m 25 $I bj")
    env.create("mgj", "lmgc")
    env.init("mgj", "$S pre post bbase $T transfer")
    env.vector("x", np.zeros(2))
    env.vector("b", np.zeros(2))
    env.create("s", "ls")
    env.init("s", "$x x $b b $A A $m 5 $I mgj $red 1e-8")
    p = env.params("s")
    assert p.base_iter == ugmg.BASE_JAC and p.base_damp[0] == 0.65
    assert p.base_maxiter == 25 and p.base_red[0] == 1e-3
    env.create("bg", "gs")
    env.init("bg", "$damp 1.0")
    env.create("gbase", "ls")
    env.init("gbase", "$red 1e-3 $m 25 $I bg")
    env.create("mgg", "lmgc")
    env.init("mgg", "$S pre post gbase $T transfer")
    env.create("s2", "ls")
    env.init("s2", "$x x $b b $A A $m 5 $I mgg $red 1e-8")
    with pytest.raises(ugmg.UgmgError):
        env.params("s2")
