"""GPU parity of ugmg_galerkin (AssembleGalerkinByMatrix,
np/algebra/transgrid.cc:1575-1700): coarse values P^T A P computed on the
device, applied from the top level down, against the oracle's restatement
(1e-12 relative); and an lmgc solve on the Galerkin hierarchy against the
oracle's solve on the same hierarchy (Jacobi defects within 1e-10)."""
import copy

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _oracle_chain(P):
    Q = copy.deepcopy(P)
    for l in range(Q.top, 0, -1):
        Q.levels[l - 1].vals = O.galerkin(O.Hierarchy(Q.levels), l)
    return Q


@pytest.mark.parametrize("which", ["jump3d", "elasticity"])
def test_device_galerkin_chain_matches_oracle(which):
    from ug_amd import problems, ugmg
    P = problems.jump3d(4, seed=3) if which == "jump3d" else problems.elasticity(3, seed=3)
    Q = _oracle_chain(P)
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    for l in range(P.top, 0, -1):
        ctx.galerkin(l)
    ctx.set_params(nu1=2, nu2=2, baselevel=0, smoother=ugmg.SMOOTH_JAC, damp=0.6, base_abslimit=1e-300)
    ctx.preprocess()
    for l in range(0, P.top):
        got = ctx.get_level_values(l)
        want = Q.levels[l].vals
        assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want)), l
    ctx.close()


def test_solve_on_galerkin_hierarchy():
    from ug_amd import problems, ugmg
    P = problems.jump3d(5, seed=4)
    Q = _oracle_chain(P)
    H = O.Hierarchy(Q.levels)
    want = O.ls_solve(H, Q.top, O.params(smoother=O.JAC, damp=0.8, baselevel=1), Q.rhs, red=1e-10, maxiter=100)
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    for l in range(P.top, 0, -1):
        ctx.galerkin(l)
    ctx.set_params(nu1=2, nu2=2, baselevel=1, smoother=ugmg.SMOOTH_JAC, damp=0.8, base_abslimit=1e-300,
                   use_graph=1)
    ctx.preprocess()
    x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=100)
    assert res.converged and res.number_of_linear_iterations == want["nit"]
    assert np.max(np.abs(d[:, 0] - want["defects"][:, 0]) / want["defects"][:, 0]) < 1e-10
    ctx.close()
