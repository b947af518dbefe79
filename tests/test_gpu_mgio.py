"""A UG multigrid file to a solve on the device: the unit cube refined red
three times (hexahedra, written as SaveMultiGrid writes it) is read
(ug_amd.mgio), each level assembled on the device from the file's elements
(ugmg_assemble_level with Q1 Laplace element matrices in the file's corner
numbering), the file's interpolation rows set as the transfer, and the
ls + lmgc Jacobi solve run; the levels equal the Q1 generator's
(problems.q1, whose transfer and stencils are pinned) in the file's node
order, and the solution equals the generator's solve mapped to that order
(Jacobi iterates depend on the order only through rounding).
"""
import numpy as np
import pytest

from test_mgio import HEX, HEX_C, HEX_E, HEX_S, Builder, hex_red_rule
import ug_mgfile as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_mg_file_to_device_solve(tmp_path):
    from ug_amd import mgio, problems, ugmg
    B = Builder(3, {HEX: (8, HEX_E, HEX_S)}, {HEX: [hex_red_rule()]}, HEX_C, [(HEX, list(range(8)))])
    for _ in range(3):
        B.refine(lambda k, t, P: 0)
    path = tmp_path / "cube.ug.mg"
    B.write(path, W.XDR)
    mg = mgio.read(path)
    Q = problems.q1(3, 1, 3, problems.KAPPA_ONE, seed=0x5EED0401)
    gperm = [c[0] + 2 * c[1] + 4 * c[2] for c in HEX_C]    # file corner k -> generator corner
    ctx = ugmg.Context(0)
    ctx.set_nlevels(len(mg["levels"]), 1)
    lexs = []
    for l, lv in enumerate(mg["levels"]):
        N = 2 ** l + 1
        ijk = np.round(lv["xyz"] * (N - 1)).astype(np.int64)
        lex = ijk[:, 0] + N * (ijk[:, 1] + N * ijk[:, 2])
        lexs.append(lex)
        G = Q.levels[l]
        el = problems.q1_elements(3, 1, l, problems.KAPPA_ONE)
        assert np.all(el.scale == 1.0)
        kref = el.kref[np.ix_(gperm, gperm)]
        ne = len(lv["elem_tag"])
        elem = lv["elem_corners"].reshape(ne, 8).astype(np.int32)
        emat = np.broadcast_to(kref, (ne, 8, 8)).copy()
        dev = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dt)
        ctx.assemble_level(l, len(lex), dev(elem, torch.int32), emat=dev(emat, torch.float64),
                           skip=dev(G.skip[lex].view(np.int32), torch.int32),
                           vclass=dev(G.vclass[lex], torch.uint8), vnclass=dev(G.vnclass[lex], torch.uint8))
        torch.cuda.synchronize()
        # the level is the generator's in the file's node order
        rp, ci, va, sk = ctx.get_level(l)
        A = np.zeros((len(lex), len(lex)))
        rows = np.repeat(np.arange(len(lex)), np.diff(rp))
        A[rows, ci] = va
        np.testing.assert_allclose(A, G.dense()[np.ix_(lex, lex)], rtol=0, atol=1e-14)
    for l in range(1, len(mg["levels"])):
        lv = mg["levels"][l]
        ctx.set_transfer(l, lv["p_ptr"], lv["p_idx"], lv["p_w"])
    ctx.set_params(nu1=2, nu2=2, gamma=1, baselevel=0, smoother=ugmg.SMOOTH_JAC, damp=0.8,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, use_graph=1)
    ctx.preprocess(len(mg["levels"]) - 1)
    top = len(mg["levels"]) - 1
    lex = lexs[-1]
    x = torch.zeros(len(lex), dtype=torch.float64, device="cuda")
    b = torch.from_numpy(Q.rhs[lex].copy()).cuda()
    torch.cuda.synchronize()
    res, d = ctx.solve(top, x, b, red=1e-10, maxiter=100)
    torch.cuda.synchronize()
    xs = x.cpu().numpy()
    ctx.close()
    ref = ugmg.Context(0)
    ref.set_hierarchy(Q)
    ref.set_params(nu1=2, nu2=2, gamma=1, baselevel=0, smoother=ugmg.SMOOTH_JAC, damp=0.8,
                   base_abslimit=1e-300, base_red=1e-12, base_maxiter=1, use_graph=1)
    ref.preprocess()
    xr, br, rr, dr = ref.solve_host(Q.top, Q.rhs, red=1e-10, maxiter=100)
    ref.close()
    assert res.converged and res.number_of_linear_iterations == rr.number_of_linear_iterations
    np.testing.assert_allclose(np.array(d)[:, 0], dr[:, 0], rtol=1e-10)
    assert np.max(np.abs(xs - xr[lex])) <= 1e-10 * np.max(np.abs(xr))
