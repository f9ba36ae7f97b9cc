"""Jacobi-preconditioned BiCGSTAB (solver option precond=3) through the C ABI.

No reference counterpart: LPM-CPP solves with MKL's CG or PARDISO
(solver.h:190-354).  The stiffness is symmetric positive definite after the
Dirichlet elimination, so the solution is unique: the GPU's BiCGSTAB solution
is compared with the oracle's (its CG is pinned against SuperLU), full
quasi-static steps with BiCGSTAB inner solves with the oracle's at the 1e-8
bar, and the iteration count with `bicgstab_np`, a plain numpy restatement of
the textbook method (Barrett et al., Templates, 1994, Fig. 2.10) on the
exported matrix.  At a size whose SpMV partials take the second reduction level,
the solution is checked by its residual.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import lpmg
from lpmg import lattice as lat
from lpmg import problems as P
from oracle import Oracle

pytestmark = pytest.mark.gpu

R = 0.1
ORC_OPTS = dict(precond=0, rtol=1e-24, atol=0.0)


def bicgstab_np(A, b, dinv, rtol, atol, maxit):
    """Preconditioned BiCGSTAB with M^-1 = diag(dinv) and r~ = r0 = b (x0 = 0);
    stops when ||r||^2 <= rtol ||r0||^2 + atol or after maxit iterations.
    Returns (x, iterations)."""
    x = np.zeros_like(b)
    r = b.copy()
    rh = b.copy()
    p = b.copy()
    y = dinv * p
    rho = rh @ r
    thresh = rtol * (r @ r) + atol
    if r @ r <= thresh:
        return x, 0
    for i in range(maxit):
        v = A @ y
        alpha = rho / (rh @ v)
        s = r - alpha * v
        x = x + alpha * y
        z = dinv * s
        t = A @ z
        tt = t @ t
        omega = (t @ s) / tt if tt > 0 else 0.0
        x = x + omega * z
        r = s - omega * t
        rho_new = rh @ r
        if r @ r <= thresh or i + 1 >= maxit:
            return x, i + 1
        beta = (rho_new / rho) * (alpha / omega)
        p = r + beta * (p - omega * v)
        y = dinv * p
        rho = rho_new
    return x, maxit


def _full_matrix(s):
    """The eliminated stiffness, both triangles, from the reference's upper CSR."""
    _, IK, JK, K = s.csr()
    n = s.dim * s.N
    U = sp.csr_matrix((K, (np.repeat(np.arange(n), np.diff(IK)), JK - 1)), shape=(n, n))
    return (U + sp.triu(U, 1).T).tocsr()


def _relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _loaded_pair(L, xyz):
    g = lpmg.System(L, R, lat.ELASTIC, xyz)
    o = Oracle(L, R, lat.ELASTIC, xyz)
    dim = o.dim
    z = xyz[:, dim - 1]
    cons = np.zeros((o.N, 3))
    cons[z < z.min() + 0.5 * R] = 1
    for s in (g, o):
        s.set_material(0, [0, 1, 200e3, 0.3, 1.0])
        s.set("CONSTRAINT", cons)
        s.set("PEX", np.tile([0.3, -0.2, 0.5], o.N) * (1 - cons.ravel()))
        s.forward()
        s.assemble()
        s.update_rr()
    return g, o


LATTICES = [("hex", lat.HEXAGON2D, lambda: lat.hex_plate(13, 11, R)),
            ("square", lat.SQUARE2D, lambda: lat.sc_block(9, 8, 1, R)),
            ("sc", lat.SC3D, lambda: lat.sc_block(7, 6, 5, R)),
            ("bcc", lat.BCC3D, lambda: lat.bcc_block(4, 4, 5, R)),
            ("fcc", lat.FCC3D, lambda: lat.fcc_block(3, 3, 4, R))]


@pytest.mark.parametrize("name,L,gen", LATTICES)
def test_bicgstab_solution_matches_oracle_and_textbook_iterations(name, L, gen):
    xyz = gen()
    g, o = _loaded_pair(L, xyz)
    A = _full_matrix(o)
    b = o.get("RESIDUAL")
    dinv = 1.0 / A.diagonal()
    # iteration count at ||r||^2 <= 1e-16 ||r0||^2: the numpy restatement on the
    # oracle's matrix takes as many, up to rounding (perturbing b by 1e-14
    # moves its count by up to 3 on these lattices; nearer the FP64 floor the
    # counts scatter more)
    st16 = g.solve_linear(precond=3, rtol=1e-16, atol=0.0)
    _, it_np = bicgstab_np(A, b, dinv, 1e-16, 0.0, 10 * len(b))
    assert abs(int(st16[0]) - it_np) <= max(3, it_np // 20), (st16[0], it_np)
    assert st16[1] <= 1e-16 * (b @ b) * (1 + 1e-12)
    # the solution at 1e-24 (solve_linear leaves the residual as it was)
    x0 = g.get("XYZ")
    st = g.solve_linear(precond=3, rtol=1e-24, atol=0.0)
    xo = o.get("XYZ")
    o.solve_linear(**ORC_OPTS)
    du_o = (o.get("XYZ") - xo).reshape(-1, 3)[:, : o.dim].ravel()
    du_g = (g.get("XYZ") - x0).reshape(-1, 3)[:, : o.dim].ravel()
    assert _relerr(du_g, du_o) < 1e-10
    x_np, _ = bicgstab_np(A, b, dinv, 1e-24, 0.0, 10 * len(b))
    assert _relerr(x_np, du_o) < 1e-10
    # the final squared residual reported is below the threshold
    assert 0 < st[0] and st[1] <= 1e-24 * (b @ b) * (1 + 1e-12)


@pytest.mark.parametrize("precision", [0, 1])
def test_bicgstab_is_deterministic(precision):
    xyz = lat.sc_block(8, 7, 6, R)
    runs = []
    for _ in range(2):
        g, _ = _loaded_pair(lat.SC3D, xyz)
        if precision:
            g.assemble(precision=1)
        x0 = g.get("XYZ")
        st = g.solve_linear(precond=3, rtol=1e-10, atol=0.0, precision=precision)
        runs.append((st.copy(), g.get("XYZ") - x0))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])


def test_bicgstab_zero_rhs_takes_no_iteration():
    xyz = lat.hex_plate(8, 8, R)
    g = lpmg.System(lat.HEXAGON2D, R, lat.ELASTIC, xyz)
    g.set_material(0, [0, 1, 200e3, 0.3, 1.0])
    g.forward()
    g.assemble()
    g.update_rr()  # nothing loaded: the residual is zero
    x0 = g.get("XYZ")
    st = g.solve_linear(precond=3, rtol=1e-12, atol=0.0)
    assert st[0] == 0
    np.testing.assert_array_equal(g.get("XYZ"), x0)


def _run_problem(prob, backend, opts, tol=1e-11):
    s, gids = P.build(backend, prob)
    s.forward()
    stats = P.solve_problem(s, gids, P.Problem(**{**prob.__dict__, "tol": tol}), opts)
    return s, stats


@pytest.mark.parametrize("prob", [P.sc_cube(n=8, steps=3), P.hex_plate(nx=14, ny=12, steps=3)],
                         ids=["sc", "hex"])
@pytest.mark.parametrize("precision", [0, 1])
def test_quasi_static_steps_with_bicgstab_match_oracle(prob, precision):
    opts = dict(precond=3, rtol=1e-20 if precision == 0 else 1e-6, atol=0.0, precision=precision)
    g, _ = _run_problem(prob, lpmg.System, opts)
    o, _ = _run_problem(prob, Oracle, ORC_OPTS)
    x0 = o.get("XYZ_INITIAL")
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8
    assert _relerr(g.get("PIN"), o.get("PIN")) < 1e-8


def test_bicgstab_with_two_level_partials_solves_the_system():
    # 281k particles: more SpMV workgroups than one level of partials holds
    # (8,192), so both SpMVs' partials take the second reduction level
    xyz = lat.hex_plate(530, 530, R)
    g = lpmg.System(lat.HEXAGON2D, R, lat.ELASTIC, xyz)
    g.set_material(0, [0, 1, 200e3, 0.3, 1.0])
    y = xyz[:, 1]
    cons = np.zeros((g.N, 3))
    cons[y < y.min() + 0.5 * R] = 1
    g.set("CONSTRAINT", cons)
    g.set("PEX", np.tile([0.3, -0.2, 0.0], g.N) * (1 - cons.ravel()))
    g.forward()
    g.assemble()
    g.update_rr()
    assert g.N > 8192 * 32
    b = g.get("RESIDUAL")
    x0 = g.get("XYZ")
    st = g.solve_linear(precond=3, rtol=1e-16, atol=0.0)
    du = (g.get("XYZ") - x0).reshape(-1, 3)[:, :2].ravel()
    res = b - g.spmv(du, 0)
    assert res @ res <= 1e-15 * (b @ b), (st, res @ res / (b @ b))
    assert 0 < st[0] < len(b)
