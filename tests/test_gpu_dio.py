"""A solve fed from and written to UG data files (include/ugmg_dio.h): the
right-hand side of the 129x129 Poisson hierarchy is stored as SaveData
would store a vector descriptor on all levels (node-major, consecutive IDs
level after level, coarse levels zero), read back, solved by ls + lmgc on
the GPU, and the solution written and re-read in each mode; the binary
modes give the stand-alone solve bit for bit."""
import numpy as np
import pytest

from ug_amd import dio, problems, ugmg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("mode", [dio.XDR, dio.BIN])
def test_solve_from_and_to_data_files(tmp_path, mode):
    P = problems.poisson2d_129()
    sizes = [lv.n for lv in P.levels]
    allv = np.concatenate([np.zeros(n) for n in sizes[:-1]] + [P.rhs])
    f_in = str(tmp_path / "rhs.ug.data")
    dio.write(f_in, allv, mode=mode, vd=[("rhs", "u")])
    h, data = dio.read(f_in)
    rhs = dio.per_level(data, 1, sizes)[-1]
    assert np.array_equal(rhs, P.rhs)
    ctx = ugmg.Context(0)
    ctx.set_hierarchy(P)
    ctx.set_params(nu1=2, nu2=2, baselevel=0, smoother=ugmg.SMOOTH_JAC, damp=0.8, base_abslimit=1e-300)
    ctx.preprocess()
    x, _, res, _ = ctx.solve_host(P.top, rhs, red=1e-10, maxiter=50)
    x_ref, _, _, _ = ctx.solve_host(P.top, P.rhs, red=1e-10, maxiter=50)
    ctx.close()
    assert res.converged and np.array_equal(x, x_ref)
    f_out = str(tmp_path / "sol.ug.data")
    dio.write(f_out, np.concatenate([np.zeros(n) for n in sizes[:-1]] + [x]), mode=mode, vd=[("sol", "u")])
    h2, back = dio.read(f_out)
    assert h2["vd"] == [("sol", 1, dio.SCALAR, "u")]
    assert np.array_equal(dio.per_level(back, 1, sizes)[-1], x)
