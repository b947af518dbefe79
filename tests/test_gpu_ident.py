"""GPU parity of the stop tests with VD_IDENT grouping (sc_cmp, np/udm/
npscan.cc:1027-1057; ugmg_set_ident) against the oracle on the 17^3 P1
elasticity system (3x3 blocks): the ls stops where the oracle stops for
every grouping, and the base solve's $abslimit test, grouped on the device,
skips the same coarse solves (defects per iteration within 1e-9 over 30
iterations of a solve that never meets its target)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _ctx(P, ident, **kw):
    from ug_amd import ugmg
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, gamma=1, smoother=ugmg.SMOOTH_JAC, damp=0.6, base_red=1e-12, base_maxiter=1, **kw)
    ctx.preprocess()
    ctx.set_ident(ident)
    return ctx


@pytest.mark.parametrize("ident", [None, [0, 0, 0], [0, 1, 1]])
def test_ls_stops_where_the_oracle_stops(elasticity_17, ident):
    P = elasticity_17
    want = O.ls_solve(O.Hierarchy(P.levels), P.top,
                      O.params(smoother=O.JAC, damp=0.6, baselevel=1, ident=ident), P.rhs, red=1e-6, maxiter=80)
    ctx = _ctx(P, ident, baselevel=1, base_abslimit=1e-300)
    try:
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-6, maxiter=80)
    finally:
        ctx.close()
    assert res.converged and res.number_of_linear_iterations == want["nit"]
    assert np.max(np.abs(d - want["defects"]) / want["defects"]) < 1e-10


def test_base_abslimit_grouped_on_the_device(elasticity_17):
    P = elasticity_17
    ident = [0, 0, 0]
    want = O.ls_solve(O.Hierarchy(P.levels), P.top,
                      O.params(smoother=O.JAC, damp=0.6, baselevel=2, base_abslimit=1e-2, ident=ident), P.rhs,
                      red=1e-12, maxiter=30)
    plain = O.ls_solve(O.Hierarchy(P.levels), P.top,
                       O.params(smoother=O.JAC, damp=0.6, baselevel=2, base_abslimit=1e-2), P.rhs,
                       red=1e-12, maxiter=30)
    assert np.max(np.abs(want["defects"] - plain["defects"]) / plain["defects"]) > 1e-3   # the grouping matters
    ctx = _ctx(P, ident, baselevel=2, base_abslimit=1e-2)
    try:
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-12, maxiter=30)
    finally:
        ctx.close()
    assert res.number_of_linear_iterations == want["nit"] == 30
    assert np.max(np.abs(d - want["defects"]) / want["defects"]) < 1e-9


def test_ident_entry_count_must_match(elasticity_17):
    from ug_amd import ugmg
    ctx = _ctx(elasticity_17, [0, 0], baselevel=1)
    try:
        with pytest.raises(ugmg.UgmgError) as e:
            ctx.solve_host(elasticity_17.top, elasticity_17.rhs, red=1e-6, maxiter=5)
        assert e.value.code == -1
        with pytest.raises(ugmg.UgmgError):
            ctx.set_ident([1, 0, 0])          # entry 1 is no group leader of itself
    finally:
        ctx.close()
