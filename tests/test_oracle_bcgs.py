"""CPU pins for the convection-diffusion workload and the BiCGStab oracle.

The P1 SUPG generator is pinned by what the discretisation fixes exactly on a
uniform triangulation (linear functions, the streamline-diffusion diagonal,
P1 interpolation of linear functions); the oracle's BCGSSolver restatement
(np/procs/ls.cc:1864-2062) is pinned against a textbook BiCGStab
(van der Vorst 1992, right-preconditioned form) written here in numpy on a
dense matrix, with the lmgc preconditioner as the dense linear operator
(I - E) A^-1 of the textbook cycle.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from ug_amd import problems

from test_oracle_pins import dense_P, mg_error_operator


def coords(N):
    h = 1.0 / (N - 1)
    iy, ix = np.divmod(np.arange(N * N), N)
    return ix * h, iy * h


def interior(N):
    iy, ix = np.divmod(np.arange(N * N), N)
    return (ix > 0) & (ix < N - 1) & (iy > 0) & (iy < N - 1)


@pytest.mark.parametrize("theta", [0.3, 2.0, 4.5])
def test_p1_cd_reproduces_linear_functions(theta):
    """-eps Lap u + b.grad u for u = x, y: the P1 Galerkin row of an interior
    node gives b_k * int phi_i = b_k h^2; diffusion and SUPG rows vanish."""
    P = problems.p1_cd(3, 1e-3, theta, seed=1)
    lv = P.levels[-1]
    N = lv.nodes_per_dir
    assert N == 17 and lv.n == N * N
    A = lv.dense()
    x, y = coords(N)
    h = 1.0 / (N - 1)
    m = interior(N)
    np.testing.assert_allclose((A @ x)[m], h * h * math.cos(theta), rtol=0, atol=1e-14)
    np.testing.assert_allclose((A @ y)[m], h * h * math.sin(theta), rtol=0, atol=1e-14)
    np.testing.assert_allclose((A @ np.ones(lv.n))[m], 0.0, atol=1e-14)
    assert abs(P.theta - theta) == 0.0


@pytest.mark.parametrize("eps", [1e-1, 1e-3])
def test_p1_cd_streamline_diffusion_diagonal(eps):
    """theta = 0: interior diagonal = eps * 4 (P1 Laplacian) + tau * 2
    (int (d_x phi_i)^2), convection diagonal 0, with
    tau = h/2 (coth Pe - 1/Pe), Pe = h / (2 eps)."""
    P = problems.p1_cd(3, eps, 0.0, seed=1)
    for lv in P.levels[1:]:
        N = lv.nodes_per_dir
        h = 1.0 / (N - 1)
        pe = h / (2 * eps)
        tau = h / 2 * (1 / math.tanh(pe) - 1 / pe)
        d = np.diag(lv.dense())[interior(N)]
        np.testing.assert_allclose(d, 4 * eps + 2 * tau, rtol=1e-13)


def test_p1_cd_dirichlet_rows_and_pattern():
    P = problems.p1_cd(3, 1e-3, 1.0, seed=1)
    for lv in P.levels:
        N = lv.nodes_per_dir
        A = lv.dense()
        bnd = ~interior(N)
        assert np.all((lv.skip == 1) == bnd)
        np.testing.assert_array_equal(A[bnd] * (1 - np.eye(lv.n))[bnd], 0.0)
        np.testing.assert_array_equal(np.diag(A)[bnd], 1.0)
        # diagonal first, 7-point pattern of the (0,0)-(1,1) split
        for v in range(lv.n):
            cols = lv.col_idx[lv.row_ptr[v]:lv.row_ptr[v + 1]]
            assert cols[0] == v
            iy, ix = divmod(v, N)
            for c in cols:
                jy, jx = divmod(int(c), N)
                assert (jx - ix, jy - iy) in {(-1, -1), (0, -1), (-1, 0), (0, 0), (1, 0), (0, 1), (1, 1)}
        rows = np.diff(lv.row_ptr)
        assert np.all(rows[interior(N)] == 7)


def test_p1_transfer_interpolates_linear_functions():
    P = problems.p1_cd(4, 1e-3, 1.0, seed=1)
    for l in range(1, P.top + 1):
        Pm = dense_P(P, l)
        xc, yc = coords(P.levels[l - 1].nodes_per_dir)
        xf, yf = coords(P.levels[l].nodes_per_dir)
        np.testing.assert_allclose(Pm @ xc, xf, atol=1e-15)
        np.testing.assert_allclose(Pm @ yc, yf, atol=1e-15)
        w = P.levels[l].p_w
        assert set(np.unique(w)) <= {0.5, 1.0}


def test_p1_cd_wind_angle_is_drawn_first():
    P = problems.p1_cd(2, 1e-3, float("nan"), seed=0x5EED0003)
    u = problems.splitmix_uniform(0x5EED0003, 1, 0.0, 2 * math.pi)[0]
    assert P.theta == u


# --------------------------------------------------------------- BiCGStab

def textbook_bicgstab(A, M, b, red, maxiter):
    """Right-preconditioned BiCGStab, shadow residual r0; norms after each
    half step (after s, after r)."""
    n = len(b)
    x = np.zeros(n)
    r = b.copy()
    rh = r.copy()
    rho = alpha = omega = 1.0
    v = np.zeros(n)
    p = np.zeros(n)
    seq = [np.linalg.norm(r)]
    target = seq[0] * red
    for _ in range(maxiter):
        rho_new = rh @ r
        beta = rho_new * alpha / (rho * omega)
        p = r + beta * (p - omega * v)
        ph = M @ p
        v = A @ ph
        alpha = rho_new / (rh @ v)
        x = x + alpha * ph
        s = r - alpha * v
        seq.append(np.linalg.norm(s))
        if seq[-1] < target:
            return x, s, np.array(seq)
        sh = M @ s
        t = A @ sh
        omega = (t @ s) / (t @ t)
        x = x + omega * sh
        r = s - omega * t
        rho = rho_new
        seq.append(np.linalg.norm(r))
        if seq[-1] < target:
            break
    return x, r, np.array(seq)


def test_bcgs_unpreconditioned_is_textbook():
    P = problems.p1_cd(2, 1e-2, 1.0, seed=5)
    H = O.Hierarchy(P.levels)
    A = P.levels[-1].dense()
    r = O.bcgs_solve(H, P.top, O.params(), P.rhs, red=1e-10, maxiter=200, use_iter=False)
    x, res, seq = textbook_bicgstab(A, np.eye(A.shape[0]), P.rhs, 1e-10, 200)
    d = r["defects"][:, 0]
    assert r["converged"] and len(d) == len(seq) == r["nit"] + 1
    np.testing.assert_allclose(d, seq, rtol=1e-6, atol=1e-11 * seq[0])
    np.testing.assert_allclose(r["x"], x, rtol=0, atol=1e-9 * np.max(np.abs(x)))
    np.testing.assert_allclose(r["b"], P.rhs - A @ r["x"], atol=1e-12 * np.max(np.abs(P.rhs)))


@pytest.mark.parametrize("kind,damp", [("jac", 0.8), ("mcgs", 1.0)])
def test_bcgs_lmgc_is_textbook_with_the_cycle_as_preconditioner(kind, damp):
    P = problems.p1_cd(3, 1e-2, 2.5, seed=6)       # 17x17, 4 levels
    H = O.Hierarchy(P.levels)
    top, bl = P.top, 0
    E = mg_error_operator(P, top, bl, kind, damp, 1, 1)
    A = P.levels[top].dense()
    M = (np.eye(A.shape[0]) - E) @ np.linalg.inv(A)
    p = O.params(smoother=O.JAC if kind == "jac" else O.MCGS, damp=damp, baselevel=bl, nu1=1, nu2=1)
    r = O.bcgs_solve(H, top, p, P.rhs, red=1e-10, maxiter=100)
    x, res, seq = textbook_bicgstab(A, M, P.rhs, 1e-10, 100)
    d = r["defects"][:, 0]
    assert r["converged"] and len(d) == len(seq)
    np.testing.assert_allclose(d, seq, rtol=1e-6, atol=1e-11 * seq[0])
    np.testing.assert_allclose(r["x"], x, atol=1e-9 * np.max(np.abs(x)))


def test_bcgs_abslimit_and_maxiter():
    P = problems.p1_cd(3, 1e-2, 2.5, seed=6)
    H = O.Hierarchy(P.levels)
    p = O.params(smoother=O.MCGS, damp=1.0, nu1=1, nu2=1)
    full = O.bcgs_solve(H, P.top, p, P.rhs, red=1e-10, maxiter=100)
    d = full["defects"][:, 0]
    cut = O.bcgs_solve(H, P.top, p, P.rhs, red=1e-10, abslimit=float(d[3]) * 1.0000001, maxiter=100)
    assert cut["converged"] and cut["nit"] == 3
    # the loop bound counts full iterations (two LinearResiduum each)
    short = O.bcgs_solve(H, P.top, p, P.rhs, red=1e-10, maxiter=2)
    assert not short["converged"] and short["nit"] == 4
