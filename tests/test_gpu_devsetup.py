"""Device-side layout build (ugmg_set_setup UGMG_SETUP_DEVICE) against the
host build, through the C ABI.

The layout is the part of LmgcPreProcess that runs over every entry of a
level: the greedy colouring, the GS device order (window chosen on samples),
the permuted level, the SELL-64 slabs of A, L, U, P and P^T (uniform slices,
shared offset lists, phantoms, padding), the diagonal blocks, classes and
colour-0 flags.  The host build is the reference here: both builds must give
every device array and every host-side description byte for byte
(ugmg_layout_digest), on scalar and 3x3-block levels, uniform and locally
refined hierarchies, GS and Jacobi smoothers; the solves that run on them
then agree bit for bit.  The device colouring alone is checked against the
host's greedy colouring on the problems' levels and on random graphs,
unsymmetric patterns and graphs needing more than 64 colours.
"""
import time
from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _ctx(P, smoother, baselevel, where, damp=1.0, **kw):
    from ug_amd import ugmg
    ctx = ugmg.Context(0)
    ctx.set_setup(where)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=baselevel, smoother=smoother, damp=damp,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, **kw)
    t = time.perf_counter()
    ctx.preprocess()
    return ctx, time.perf_counter() - t


def _csr(n, pairs):
    """CSR with the diagonal first, then the listed (row, col) entries."""
    rows = [[i] for i in range(n)]
    for r, c in pairs:
        if r != c and c not in rows[r]:
            rows[r].append(c)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([c for r in rows for c in r], np.int32)
    return SimpleNamespace(n=n, row_ptr=rp, col_idx=ci)


def _graphs():
    rng = np.random.default_rng(0x5EED0101)
    out = {}
    n = 3000   # symmetric random graph
    e = rng.integers(0, n, size=(12000, 2))
    out["random_sym"] = _csr(n, [(a, b) for a, b in e] + [(b, a) for a, b in e])
    # unsymmetric pattern: the transposed lower entries matter
    out["random_unsym"] = _csr(n, [(a, b) for a, b in rng.integers(0, n, size=(15000, 2))])
    # a 100-clique plus a path: 100 colours (the >64 branch)
    k = 100
    out["clique100"] = _csr(k + 50, [(a, b) for a in range(k) for b in range(k)]
                            + [(k + i, k + i + 1) for i in range(49)] + [(k + i + 1, k + i) for i in range(49)])
    out["path_reversed"] = _csr(500, [(i + 1, i) for i in range(499)])   # lower entries only
    return out


@pytest.mark.parametrize("name", ["random_sym", "random_unsym", "clique100", "path_reversed"])
def test_color_device_random_graphs(name):
    from ug_amd import ugmg
    g = _graphs()[name]
    nh, ch = ugmg.color(g)
    ctx = ugmg.Context(0)
    try:
        nd, cd, rounds = ctx.color_device(g)
    finally:
        ctx.close()
    print(name, "colours", nh, "rounds", rounds)
    assert nd == nh and np.array_equal(cd, ch)
    if name == "clique100":
        assert nh == 100


def test_color_device_problem_levels(jump3d_33, poisson2d):
    from ug_amd import problems, ugmg
    ctx = ugmg.Context(0)
    try:
        for P in (jump3d_33, poisson2d, problems.elasticity(3), problems.p1_local(4, 2)):
            for lv in P.levels[1:]:
                nh, ch = ugmg.color(lv)
                nd, cd, rounds = ctx.color_device(lv)
                print(P.name, lv.n, "colours", nh, "rounds", rounds)
                assert nd == nh and np.array_equal(cd, ch)
    finally:
        ctx.close()


def _compare_builds(P, smoother, baselevel, damp=1.0, solve=True):
    from ug_amd import ugmg
    ch, th = _ctx(P, smoother, baselevel, ugmg.SETUP_HOST, damp)
    cd, td = _ctx(P, smoother, baselevel, ugmg.SETUP_DEVICE, damp)
    try:
        bl = ch.base_level()
        assert cd.base_level() == bl
        for l in range(bl, P.top + 1):
            dh, dd = ch.layout_digest(l), cd.layout_digest(l)
            bad = np.nonzero(dh != dd)[0]
            assert bad.size == 0, (P.name, l, bad.tolist())
        if solve:
            n = P.levels[-1].n * P.levels[-1].b
            rhs = np.asarray(P.rhs, np.float64) if getattr(P, "rhs", None) is not None \
                else np.random.default_rng(3).uniform(-1, 1, n)
            res = []
            for c in (ch, cd):
                x = torch.zeros(n, dtype=torch.float64, device="cuda")
                b = torch.from_numpy(rhs.copy()).cuda()
                torch.cuda.synchronize()
                r, defects = c.solve(P.top, x, b, red=1e-8, maxiter=60)
                torch.cuda.synchronize()
                res.append((x.cpu().numpy(), np.asarray(defects)))
            assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    finally:
        ch.close()
        cd.close()
    return th, td


def test_layout_jump3d_33_gs(jump3d_33):
    from ug_amd import ugmg
    th, td = _compare_builds(jump3d_33, ugmg.SMOOTH_GS, 2)
    print(f"33^3 GS: host build {th:.3f} s, device build {td:.3f} s")


def test_layout_jump3d_33_jacobi(jump3d_33):
    from ug_amd import ugmg
    _compare_builds(jump3d_33, ugmg.SMOOTH_JAC, 2, damp=0.8)


def test_layout_poisson2d_gs(poisson2d):
    from ug_amd import ugmg
    _compare_builds(poisson2d, ugmg.SMOOTH_GS, 0)


def test_layout_elasticity_blocks_gs():
    from ug_amd import problems, ugmg
    _compare_builds(problems.elasticity(4), ugmg.SMOOTH_GS, 1)


def test_layout_locally_refined():
    """Red-green hierarchies: classes below 3 on some rows (U keeps every
    off-diagonal entry of an inactive row, L drops columns of inactive
    vectors, P^T keeps only VCLASS >= 2 fine rows)."""
    from ug_amd import problems, ugmg
    for P in (problems.p1_local(4, 3), problems.elasticity_local(3, 2)):
        _compare_builds(P, ugmg.SMOOTH_GS, 1, solve=False)
        _compare_builds(P, ugmg.SMOOTH_JAC, 1, damp=0.8, solve=False)


def test_layout_convection_diffusion():
    from ug_amd import problems, ugmg
    P = problems.p1_cd(5, 1e-3, float("nan"), seed=0x5EED0003)
    _compare_builds(P, ugmg.SMOOTH_GS, 1, solve=False)


def test_layout_jump3d_129_full_size():
    """The bench's 129^3 hierarchy: both builds equal on every level, and the
    device build's wall time beside the host build's."""
    from ug_amd import problems, ugmg
    P = problems.jump3d(7)
    th, td = _compare_builds(P, ugmg.SMOOTH_GS, 4, solve=False)
    print(f"129^3 GS: host build {th:.3f} s, device build {td:.3f} s")


def test_assembled_level_laid_out_from_its_device_csr(jump3d_33):
    """ugmg_assemble_level keeps the flattened level on the device and the
    device layout build reads it there (no upload): the layout equals the
    one built from the generator's host level, byte for byte."""
    from ug_amd import problems, ugmg
    P = jump3d_33
    el = problems.q1_elements(3, 1, 5, problems.KAPPA_JUMP, seed=0x5EED0002)
    elem = torch.from_numpy(el.elem).cuda()
    kref = torch.from_numpy(el.kref).cuda()
    scale = torch.from_numpy(el.scale).cuda()
    skip = torch.from_numpy(el.skip.view(np.int32)).cuda()
    torch.cuda.synchronize()
    ref, _ = _ctx(P, ugmg.SMOOTH_GS, 2, ugmg.SETUP_HOST)
    ctx = ugmg.Context(0)
    try:
        ctx.set_hierarchy(P)
        ctx.assemble_level(P.top, el.n, elem, kref=kref, scale=scale, skip=skip)
        ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=2, smoother=ugmg.SMOOTH_GS, damp=1.0,
                       base_abslimit=1e-300, base_red=1e-12, base_maxiter=1)
        ctx.preprocess()
        for l in range(2, P.top + 1):
            assert np.array_equal(ctx.layout_digest(l), ref.layout_digest(l)), l
        ctx.preprocess()   # again: the kept CSR was consumed, the host copy serves
        assert np.array_equal(ctx.layout_digest(P.top), ref.layout_digest(P.top))
    finally:
        ctx.close()
        ref.close()
