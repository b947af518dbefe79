"""Device-side assembly (ugmg_assemble_level): element lists and element
matrices in device memory become the level's flattened matrix on the GPU
-- the element loop of AddVlistMValues / AddVlistVValues
(np/udm/disctools.cc:536-555, 841-950) and AssembleDirichletBoundary
(disctools.cc:1837-1875).  The level it sets is bit-identical to the
generators' finest level (which element-order summation on the host
reproduces, tests/test_assemble_host.py), a solve on it is bit-identical to
the solve on the host-built level, the load vector and Dirichlet values
match the host's element-order sums, and bad element lists are refused
before any index is used."""
import numpy as np
import pytest

from ug_amd import problems, ugmg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a, dtype=None):
    if a is None:
        return None
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def _assemble(ctx, level, el, evec=None, dirichlet=None, rhs=None):
    ctx.assemble_level(level, el.n, _dev(el.elem), emat=_dev(el.emat), kref=_dev(el.kref),
                       scale=_dev(el.scale), evec=_dev(evec), skip=_dev(el.skip.view(np.int32)),
                       dirichlet=_dev(dirichlet), rhs=rhs)
    ctx.sync()


CASES = {
    "q1_3d_jump": (lambda: problems.q1(3, 1, 4, problems.KAPPA_JUMP, seed=31),
                   lambda: problems.q1_elements(3, 1, 4, problems.KAPPA_JUMP, seed=31)),
    "q1_2d": (lambda: problems.q1(2, 2, 5, problems.KAPPA_ONE, seed=32),
              lambda: problems.q1_elements(2, 2, 5, problems.KAPPA_ONE, seed=32)),
    "p1_cd": (lambda: problems.p1_cd(4, eps=1e-3, seed=33),
              lambda: problems.p1_cd_elements(4, eps=1e-3, seed=33)),
    "elasticity": (lambda: problems.elasticity(3, seed=34),
                   lambda: problems.elasticity_elements(3, seed=34)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_assembly_is_the_generator_level(name):
    gen, elements = CASES[name]
    P = gen()
    el = elements()
    top = P.levels[-1]
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    _assemble(ctx, P.top, el)
    rp, ci, vals, skip = ctx.get_level(P.top)
    ctx.close()
    assert np.array_equal(rp, top.row_ptr)
    assert np.array_equal(ci, top.col_idx)
    assert np.array_equal(vals.view(np.int64), top.vals.view(np.int64))
    assert np.array_equal(skip, top.skip)


def test_solve_on_device_assembled_level_is_the_host_solve():
    P = problems.q1(3, 1, 5, problems.KAPPA_JUMP, seed=35)
    el = problems.q1_elements(3, 1, 5, problems.KAPPA_JUMP, seed=35)
    prm = dict(nu1=2, nu2=2, baselevel=2, smoother=ugmg.SMOOTH_GS, damp=1.0, base_abslimit=1e-300)
    out = []
    for device_level in (False, True):
        ctx = ugmg.Context(0)
        ctx.set_hierarchy(P)
        if device_level:
            _assemble(ctx, P.top, el)
        ctx.set_params(**prm)
        ctx.preprocess()
        x, b, res, d = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=100)
        ctx.close()
        out.append((x, d, res.number_of_linear_iterations))
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("which", ["q1_2d", "elasticity"])
def test_device_load_vector_and_dirichlet_values(which):
    el = problems.q1_elements(2, 1, 3) if which == "q1_2d" else problems.elasticity_elements(2)
    b = el.b
    rng = np.random.default_rng(7)
    evec = rng.uniform(-1, 1, (el.nelem, el.nc, b))
    dval = rng.uniform(-1, 1, el.n * b)
    want = problems.assemble_host(el, evec=evec, dirichlet=dval)
    ctx = ugmg.Context(0)
    ctx.set_nlevels(1, b)
    rhs = torch.zeros(el.n * b, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _assemble(ctx, 0, el, evec=evec, dirichlet=dval, rhs=rhs)
    rp, ci, vals, skip = ctx.get_level(0)
    ctx.close()
    assert np.array_equal(rp, want[0]) and np.array_equal(ci, want[1])
    assert np.array_equal(vals, want[2])
    assert np.array_equal(rhs.cpu().numpy(), want[3])


def test_bad_element_lists_are_refused():
    el = problems.q1_elements(2, 1, 2)
    ctx = ugmg.Context(0)
    ctx.set_nlevels(1, 1)
    bad = el.elem.copy()
    bad[3, 1] = el.n          # one corner outside [0, n)
    with pytest.raises(ugmg.UgmgError) as e:
        ctx.assemble_level(0, el.n, _dev(bad), kref=_dev(el.kref), scale=_dev(el.scale))
    assert e.value.code == -1 and "out of range" in str(e.value)
    # 100 elements meeting at vector 0
    many = np.tile(np.array([[0, 1, 2, 3]], np.int32), (100, 1))
    with pytest.raises(ugmg.UgmgError) as e:
        ctx.assemble_level(0, 4, _dev(many), kref=_dev(el.kref), scale=_dev(np.ones(100)))
    assert e.value.code == -1 and "96" in str(e.value)
    # the context still works afterwards
    ctx.assemble_level(0, el.n, _dev(el.elem), kref=_dev(el.kref), scale=_dev(el.scale))
    ctx.close()
