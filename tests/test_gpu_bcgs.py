"""GPU parity of the batched BCGSSolver and the batched lmgc cycle.

libugmg's BiCGStab (np/procs/ls.cc:1864-2062) through its C ABI against the
oracle's restatement, on the SUPG P1 convection-diffusion hierarchy (65x65
finest level).  Bars: per-half-step defect norms within 1e-6 relative, or
absolutely within 5e-11 of the first defect for the smallest ones
(BiCGStab amplifies the rounding differences of FMA contraction and
tree-ordered dot products); solution within 1e-8 relative; the columns of a
batched solve bit-identical to single solves.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def cd5():
    from ug_amd import problems
    return problems.p1_cd(5, 1e-2, float("nan"), seed=0x5EED0003)


def _ctx(P, smoother, damp, nrhs=1, precision=0, baselevel=0, nu=1, graph=1):
    from ug_amd import ugmg
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=nu, nu2=nu, gamma=1, baselevel=baselevel, smoother=smoother, damp=damp,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, precision=precision,
                   use_graph=graph)
    ctx.set_batch(nrhs)
    ctx.preprocess()
    return ctx


def _close(gpu, orc):
    first = orc[0]
    np.testing.assert_allclose(gpu, orc, rtol=1e-6, atol=5e-11 * first)


@pytest.mark.parametrize("kind", ["mcgs", "jac", "none"])
def test_bcgs_matches_oracle(cd5, kind):
    from ug_amd import ugmg
    P = cd5
    sm = ugmg.SMOOTH_JAC if kind == "jac" else ugmg.SMOOTH_GS
    damp = 0.8 if kind == "jac" else 1.0
    ctx = _ctx(P, sm, damp)
    use_iter = kind != "none"
    print(f"bcgs {kind} n={P.levels[-1].n}", flush=True)
    x, b, res, seqs = ctx.bcgs_host(P.top, P.rhs[None, :], red=1e-10, maxiter=150, use_iter=use_iter)
    H = O.Hierarchy(P.levels)
    p = O.params(smoother=O.JAC if kind == "jac" else O.MCGS, damp=damp, baselevel=0, nu1=1, nu2=1)
    ref = O.bcgs_solve(H, P.top, p, P.rhs, red=1e-10, maxiter=150, use_iter=use_iter)
    assert res[0].converged and ref["converged"]
    assert res[0].number_of_linear_iterations == ref["nit"]
    _close(seqs[0], ref["defects"][:, 0])
    assert np.max(np.abs(x[0] - ref["x"])) < 1e-8 * np.max(np.abs(ref["x"]))
    A = P.levels[-1].dense()
    assert np.max(np.abs(b[0] - (P.rhs - A @ x[0]))) < 1e-10 * np.max(np.abs(P.rhs))
    ctx.close()


def test_bcgs_restart_matches_oracle(cd5):
    from ug_amd import ugmg
    P = cd5
    ctx = _ctx(P, ugmg.SMOOTH_GS, 1.0)
    x, b, res, seqs = ctx.bcgs_host(P.top, P.rhs[None, :], red=1e-10, maxiter=150, restart=3)
    H = O.Hierarchy(P.levels)
    ref = O.bcgs_solve(H, P.top, O.params(smoother=O.MCGS, damp=1.0, nu1=1, nu2=1), P.rhs, red=1e-10,
                       maxiter=150, restart=3)
    assert res[0].number_of_linear_iterations == ref["nit"]
    _close(seqs[0], ref["defects"][:, 0])
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_bcgs_batched_columns_equal_single_solves(cd5, precision):
    """Each column of a batched solve (own convergence, masked once done)
    is bit-identical to the same right-hand side solved alone."""
    from ug_amd import ugmg
    P = cd5
    n = P.levels[-1].n
    rng = np.random.default_rng(11)
    skip = P.levels[-1].skip == 1
    cols = [P.rhs, rng.uniform(-1, 1, n), 1e3 * rng.uniform(-1, 1, n), np.zeros(n), rng.uniform(0, 1, n)]
    R = np.array(cols)
    R[:, skip] = 0.0
    ctx = _ctx(P, ugmg.SMOOTH_GS, 1.0, nrhs=len(cols), precision=precision)
    print(f"batched bcgs precision={precision} nrhs={len(cols)}", flush=True)
    xb, bb, resb, seqb = ctx.bcgs_host(P.top, R, red=1e-8, maxiter=100)
    for j in range(len(cols)):
        x1, b1, r1, s1 = ctx.bcgs_host(P.top, R[j:j + 1], red=1e-8, maxiter=100)
        assert r1[0].number_of_linear_iterations == resb[j].number_of_linear_iterations
        assert r1[0].converged == resb[j].converged
        assert np.array_equal(x1[0], xb[j]) and np.array_equal(b1[0], bb[j])
        assert np.array_equal(s1[0], seqb[j])
    assert resb[3].converged and resb[3].number_of_linear_iterations == 0   # zero rhs
    for j in (0, 1, 2, 4):
        assert resb[j].converged
        assert resb[j].last_defect[0] < 1e-8 * resb[j].first_defect[0]
    ctx.close()


def test_bcgs_mixed_precision_reaches_fp64_accuracy(cd5):
    from ug_amd import ugmg
    P = cd5
    A = P.levels[-1].dense()
    ctx = _ctx(P, ugmg.SMOOTH_GS, 1.0, precision=ugmg.MIXED)
    x, b, res, seqs = ctx.bcgs_host(P.top, P.rhs[None, :], red=1e-10, maxiter=150)
    assert res[0].converged
    true = np.linalg.norm(P.rhs - A @ x[0])
    assert true < 2e-10 * np.linalg.norm(P.rhs)
    xs = np.linalg.solve(A, P.rhs)
    assert np.max(np.abs(x[0] - xs)) < 1e-7 * np.max(np.abs(xs))
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_iter_batch_columns_equal_single_iter(cd5, precision):
    from ug_amd import ugmg
    P = cd5
    n = P.levels[-1].n
    rng = np.random.default_rng(5)
    R = rng.uniform(-1, 1, (3, n))
    R[:, P.levels[-1].skip == 1] = 0.0
    ld = n + 7
    ctx = _ctx(P, ugmg.SMOOTH_GS, 1.0, nrhs=3, precision=precision, graph=0)
    Bt = torch.zeros((3, ld), dtype=torch.float64, device="cuda")
    Bt[:, :n] = torch.from_numpy(R).cuda()
    Ct = torch.zeros_like(Bt)
    torch.cuda.synchronize()
    ctx.iter_batch(P.top, Ct, Bt)
    ctx.sync()
    Cb, Bb = Ct.cpu().numpy()[:, :n], Bt.cpu().numpy()[:, :n]
    for j in range(3):
        c1, b1 = ctx.iter_host(P.top, R[j])
        assert np.array_equal(c1, Cb[j]) and np.array_equal(b1, Bb[j])
    # and the FP64 cycle is the oracle's
    if precision == 0:
        H = O.Hierarchy(P.levels)
        c0, b0 = O.lmgc_iter(H, P.top, O.params(smoother=O.MCGS, damp=1.0, nu1=1, nu2=1), R[0])
        assert np.max(np.abs(Cb[0] - c0)) < 1e-12 * np.max(np.abs(c0))
    ctx.close()
