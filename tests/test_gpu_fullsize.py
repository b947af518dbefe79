"""GPU checks at the workloads' full sizes, where the oracle would take
minutes, through properties that do not depend on size:

* 129^3 jump-coefficient Q1 (2,146,689 unknowns, 57M entries): the device
  layout decompresses to UG's pattern bit for bit; dmatmul_minus equals a
  CSR product on the host within FP64 rounding; the FP64 and the mixed
  lmgc solves both reach |f - A x| / |f| <= 2e-10 (true defect, recomputed
  on the host), agree with each other to 1e-8, and every step contracts the
  defect (rate below 0.8 per FP64 cycle, 0.5 per mixed step of 3 FP32
  cycles); with adaptive mixed steps (the last step's cycles predicted from
  the rate) likewise, in no more steps; the device assembly of
  the finest level equals the host level.
* 513^2 SUPG convection-diffusion, BiCGStab + V(1,1) MC-GS with 32 columns:
  every column's true defect <= 1.1e-8 relative, and a column solved on its
  own is bit-identical to the same column inside the batch.
* 65^3 P1 elasticity (3x3 blocks): FP64 and mixed solves reach 1e-10 per
  component (true defect per component, recomputed on the host).
* Against the oracle at full size: the FP64 and mixed MC-GS solutions of
  the 129^3 problem within 1e-8 relative of the oracle's lexicographic-GS
  solve to the same tolerance (UG's l_lgs order), and two of the 32 FP64
  BiCGStab columns at 513^2 with the oracle's half-step defects (1e-6
  relative, or 5e-11 of the first) and x within 1e-8.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
from ug_amd import problems, ugmg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _csr(lv):
    # copies: scipy sorts shared index arrays in place (the level keeps UG's
    # diagonal-first rows)
    if lv.b == 1:
        return sp.csr_matrix((lv.vals.copy(), lv.col_idx.copy(), lv.row_ptr.copy()), shape=(lv.n, lv.n))
    return sp.bsr_matrix((lv.vals.reshape(-1, lv.b, lv.b).copy(), lv.col_idx.copy(), lv.row_ptr.copy()),
                         shape=(lv.n * lv.b, lv.n * lv.b)).tocsr()


@pytest.fixture(scope="module")
def jump129():
    return problems.jump3d(7)


def _ctx(P, precision, **kw):
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    prm = dict(nu1=2, nu2=2, gamma=1, baselevel=4, smoother=ugmg.SMOOTH_GS, damp=1.0,
               base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, use_graph=1)
    prm.update(kw)
    ctx.set_params(**prm, precision=precision, inner_cycles=3)
    ctx.preprocess()
    return ctx


def test_129_layout_and_spmv(jump129):
    P = jump129
    top = P.levels[-1]
    ctx = _ctx(P, ugmg.FP64)
    rp, ci, va = ctx.pattern(P.top)
    assert np.array_equal(rp, top.row_ptr) and np.array_equal(ci, top.col_idx)
    assert np.array_equal(va, top.vals)
    rng = np.random.default_rng(129)
    x = rng.uniform(-1, 1, top.n)
    y = rng.uniform(-1, 1, top.n)
    A = _csr(top)
    want = y - A @ x
    xd = torch.from_numpy(x).cuda()
    yd = torch.from_numpy(y).cuda()
    torch.cuda.synchronize()
    ctx.dmatmul_minus(P.top, yd, xd)          # dmatmul_minus(x, M, y): x -= M y
    ctx.sync()
    got = yd.cpu().numpy()
    ctx.close()
    scale = np.abs(A).max() * np.abs(x).max() * 27
    assert np.max(np.abs(got - want)) <= 1e-14 * scale


def test_129_fp64_and_mixed_solves(jump129):
    P = jump129
    A = _csr(P.levels[-1])
    f = P.rhs
    out = {}
    steps = {}
    for name, prec, adapt in (("fp64", ugmg.FP64, 1), ("mixed", ugmg.MIXED, 0), ("adaptive", ugmg.MIXED, 1)):
        ctx = _ctx(P, prec, inner_adapt=adapt)
        x, b, res, d = ctx.solve_host(P.top, f, red=1e-10, maxiter=200)
        ctx.close()
        assert res.converged
        true = np.linalg.norm(f - A @ x) / np.linalg.norm(f)
        assert true <= 2e-10, (name, true)
        d = np.array(d)[:, 0]
        rates = d[1:] / d[:-1]      # every step contracts (so the defects decrease)
        # fixed mixed steps run 3 FP32 cycles each; an adaptive last step may run one
        bound = {"fp64": 0.8, "mixed": 0.5, "adaptive": 0.5}[name]
        assert rates[:-1].max(initial=0.0) < bound and rates[-1] < 0.8, (name, rates)
        out[name] = x
        steps[name] = res.number_of_linear_iterations
    assert steps["adaptive"] <= steps["mixed"]
    for name in ("mixed", "adaptive"):
        rel = np.linalg.norm(out[name] - out["fp64"]) / np.linalg.norm(out["fp64"])
        assert rel <= 1e-8, name


def test_129_device_assembly_is_the_host_level(jump129):
    P = jump129
    top = P.levels[-1]
    el = problems.q1_elements(3, 1, 7, problems.KAPPA_JUMP, seed=0x5EED0002)
    ctx = ugmg.Context(0)
    ctx.set_nlevels(1, 1)
    ctx.assemble_level(0, el.n, torch.from_numpy(el.elem).cuda(), kref=torch.from_numpy(el.kref).cuda(),
                       scale=torch.from_numpy(el.scale).cuda(),
                       skip=torch.from_numpy(el.skip.view(np.int32)).cuda())
    rp, ci, va, sk = ctx.get_level(0)
    ctx.close()
    assert np.array_equal(rp, top.row_ptr) and np.array_equal(ci, top.col_idx)
    assert np.array_equal(va.view(np.int64), top.vals.view(np.int64))
    assert np.array_equal(sk, top.skip)


def test_513_bcgs_batch_columns():
    C = problems.p1_cd(8, seed=0x5EED0003)
    A = _csr(C.levels[-1])
    n = C.levels[-1].n
    bl = max(l for l in range(C.top + 1) if C.levels[l].n <= 729)
    rng = np.random.default_rng(513)
    R = rng.uniform(-1, 1, (32, n))
    R[:, C.levels[-1].skip.astype(bool)] = 0.0
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(C)
    ctx.set_params(nu1=1, nu2=1, gamma=1, baselevel=bl, smoother=ugmg.SMOOTH_GS, damp=1.0,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, use_graph=1,
                   precision=ugmg.MIXED, inner_cycles=1)
    ctx.set_batch(32)
    ctx.preprocess()
    X, B, results, _ = ctx.bcgs_host(C.top, R, red=1e-8, maxiter=200)
    for j in range(32):
        assert results[j].converged
        true = np.linalg.norm(R[j] - A @ X[j]) / np.linalg.norm(R[j])
        assert true <= 1.1e-8, (j, true)
    x1, _, r1, _ = ctx.bcgs_host(C.top, R[5:6], red=1e-8, maxiter=200)
    ctx.close()
    assert r1[0].number_of_linear_iterations == results[5].number_of_linear_iterations
    assert np.array_equal(x1[0], X[5])


def test_65_elasticity_solves():
    P = problems.elasticity(6)
    A = _csr(P.levels[-1])
    f = P.rhs
    bl = max(l for l in range(P.top + 1) if P.levels[l].n * 3 <= 8192)
    for prec in (ugmg.FP64, ugmg.MIXED):
        ctx = _ctx(P, prec, baselevel=bl)
        x, b, res, d = ctx.solve_host(P.top, f, red=1e-10, maxiter=200)
        ctx.close()
        assert res.converged
        r = f - A @ x
        for c in range(3):
            assert np.linalg.norm(r[c::3]) <= 2e-10 * np.linalg.norm(f[c::3]), (prec, c)


def test_129_mcgs_solution_matches_oracle_lexicographic_gs(jump129):
    P = jump129
    print("129^3 oracle lexicographic GS solve (about half a minute on the host)", flush=True)
    H = O.Hierarchy(P.levels)
    ref = O.ls_solve(H, P.top, O.params(smoother=O.GS, damp=1.0, baselevel=4), P.rhs, red=1e-10, maxiter=500)
    assert ref["converged"]
    for prec in (ugmg.FP64, ugmg.MIXED):
        ctx = _ctx(P, prec)
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=200)
        ctx.close()
        assert res.converged
        rel = np.linalg.norm(x - ref["x"]) / np.linalg.norm(ref["x"])
        print(f"precision {prec}: GPU MC-GS vs oracle lexicographic GS {rel:.2e} "
              f"({res.number_of_linear_iterations} vs {ref['nit']} iterations)")
        assert rel <= 1e-8, (prec, rel)


def test_513_bcgs_columns_match_oracle():
    C = problems.p1_cd(8, seed=0x5EED0003)
    n = C.levels[-1].n
    bl = max(l for l in range(C.top + 1) if C.levels[l].n <= 729)
    rng = np.random.default_rng(513)
    R = rng.uniform(-1, 1, (32, n))
    R[:, C.levels[-1].skip.astype(bool)] = 0.0
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(C)
    ctx.set_params(nu1=1, nu2=1, gamma=1, baselevel=bl, smoother=ugmg.SMOOTH_GS, damp=1.0,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, use_graph=1, precision=ugmg.FP64)
    ctx.set_batch(32)
    ctx.preprocess()
    X, B, results, seqs = ctx.bcgs_host(C.top, R, red=1e-8, maxiter=200)
    ctx.close()
    H = O.Hierarchy(C.levels)
    p = O.params(smoother=O.MCGS, damp=1.0, baselevel=bl, nu1=1, nu2=1)
    for j in (0, 17):
        ref = O.bcgs_solve(H, C.top, p, R[j], red=1e-8, maxiter=200)
        assert ref["converged"] and results[j].converged
        assert results[j].number_of_linear_iterations == ref["nit"], j
        orc = ref["defects"][:, 0]
        gpu = np.asarray(seqs[j])[:len(orc)]
        np.testing.assert_allclose(gpu, orc, rtol=1e-6, atol=5e-11 * orc[0])
        rel = np.linalg.norm(X[j] - ref["x"]) / np.linalg.norm(ref["x"])
        assert rel <= 1e-8, (j, rel)
