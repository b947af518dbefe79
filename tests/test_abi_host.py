"""C ABI and host-side logic of libugmg without a GPU."""
import ctypes as C
import re

import numpy as np
import pytest

from oracle import oracle as O
from ug_amd import problems, ugmg


def test_libugmg_exports_every_declared_symbol():
    lib = C.CDLL(ugmg.LIB_PATH)
    names = ugmg.declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("header", ["ugmg_np.h", "ugmg_dio.h", "ugmg_mgio.h"])
def test_libugmg_exports_the_numproc_and_data_file_symbols(header):
    names = ugmg.declared_symbols(ugmg.HEADER.replace("ugmg.h", header))
    assert names
    lib = C.CDLL(ugmg.LIB_PATH)
    assert [n for n in names if not hasattr(lib, n)] == []


def test_problem_library_exports_every_declared_symbol():
    text = open(ugmg.HEADER.replace("ugmg.h", "ugmg_problems.h")).read()
    names = sorted(set(re.findall(r"\b(ugmg_[a-z0-9_]+)\s*\(", text)))
    lib = C.CDLL(problems.LIB_PATH)
    assert [n for n in names if not hasattr(lib, n)] == []


def test_create_without_gpu_reports_nodev():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    lib = ugmg.load()
    h = C.c_void_p()
    assert lib.ugmg_create(0, C.byref(h)) == -7       # UGMG_E_NODEV
    assert lib.ugmg_solve(None, 0, None, None, None, None, 1, 1, None, ugmg.DEFECT_CB(), None) == -1


def test_params_default_are_numproc_defaults():
    p = ugmg.default_params()
    assert (p.nu1, p.nu2, p.gamma, p.baselevel) == (1, 1, 1, 0)
    assert p.base_abslimit[0] == 1e-10 and p.base_red[0] == 1e-12 and p.base_maxiter == 1
    assert p.damp[0] == 1.0 and p.ex_reorder == 1


def test_block_limit_matches_the_headers():
    hdr = open(ugmg.HEADER).read()
    assert int(re.search(r"#define UGMG_MAXB (\d+)", hdr).group(1)) == ugmg.MAXB
    ohdr = open(O.LIB_PATH.replace("liboracle.so", "ugmg_oracle.h")).read()
    assert int(re.search(r"#define ORC_MAXB (\d+)", ohdr).group(1)) == O.MAXB
    # every component written by ugmg_params_default, the fields after the arrays in place
    p = ugmg.default_params()
    assert list(p.damp) == [1.0] * ugmg.MAXB and list(p.base_red) == [1e-12] * ugmg.MAXB
    assert (p.base_maxiter, p.ex_reorder, p.precision, p.inner_cycles, p.inner_adapt) == (1, 1, 0, 1, 1)
    p = ugmg.default_params(damp=[0.5, 0.6])
    assert list(p.damp) == [0.5] + [0.6] * (ugmg.MAXB - 1)


def test_host_colouring_equals_oracle(poisson2d, jump3d_17):
    for P in (poisson2d, jump3d_17):
        H = O.Hierarchy(P.levels)
        for l in range(P.top + 1):
            nc, col = ugmg.color(P.levels[l])
            nw, want = O.greedy_color(H, l)
            assert nc == nw and np.array_equal(col, want)


def test_host_coarse_inverse_is_ex_band_lu(jump3d_17):
    H = O.Hierarchy(jump3d_17.levels)
    for l in (1, 2, 3):
        lv = jump3d_17.levels[l]
        rc, inv, order, bw = ugmg.coarse_inverse(lv, reorder=1)
        assert rc == 0
        ex = O.Ex(H, l, reorder=1)
        assert np.array_equal(order, ex.order) and bw == ex.bandwidth
        assert np.max(np.abs(inv @ lv.dense() - np.eye(lv.n))) < 1e-12


def test_host_coarse_inverse_zero_pivot():
    P = problems.q1(2, 2, 1, problems.KAPPA_ONE, seed=1)
    lv = P.levels[1]
    lv.vals = lv.vals.copy()
    lv.vals[lv.row_ptr[0]] = 0.0
    rc, *_ = ugmg.coarse_inverse(lv, reorder=0)
    assert rc == -5
