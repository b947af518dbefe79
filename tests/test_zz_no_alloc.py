"""ugmg_dnrm2x_noalloc, the entry point that allocates nothing, against
ugmg_dnrm2x: the same norms bit for bit on two calls in a row, the second
on a smaller level, for scalar and 3x3-block hierarchies."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _vec(n, seed):
    a = np.random.default_rng(seed).uniform(-1.0, 1.0, n)
    t = torch.from_numpy(a).to("cuda")
    torch.cuda.synchronize()   # the library runs on its own stream
    return a, t


@pytest.mark.parametrize("kind", ["scalar_65^3", "blocks_17^3"])
def test_dnrm2x_noalloc_equals_dnrm2x(kind):
    from ug_amd import problems, ugmg
    if kind.startswith("scalar"):
        P = problems.jump3d(6)     # 274,625 rows: the reduction's 1024-block cap
        levels = (P.top, P.top - 2)
    else:
        P = problems.elasticity(4)  # 4913 nodes x 3 components
        levels = (P.top, P.top - 1)
    print(f"case {kind}: levels {levels}", flush=True)
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=1, nu2=1, gamma=1, baselevel=1, smoother=ugmg.SMOOTH_JAC, damp=0.8,
                   base_abslimit=1e-300)
    ctx.preprocess()
    try:
        for k, l in enumerate(levels):
            lv = P.levels[l]
            a, x = _vec(lv.n * lv.b, 100 + k)
            want = np.atleast_1d(ctx.dnrm2x(l, x))
            got = np.atleast_1d(ctx.dnrm2x_noalloc(l, x))
            assert np.array_equal(got.view(np.int64), want.view(np.int64)), (l, got, want)
            exact = np.sqrt((a.reshape(-1, lv.b) ** 2).sum(axis=0))
            assert np.max(np.abs(got - exact) / exact) < 1e-13
    finally:
        ctx.close()
