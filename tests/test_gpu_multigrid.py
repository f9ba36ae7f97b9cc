"""Multigrid-preconditioned CG (opts precond = 2) on the GPU.

The preconditioner changes the CG's path, not the system it solves, so the
quasi-static results must meet the same bar as the Jacobi CG: displacements and
forces within 1e-8 relative L2 of the CPU oracle at a tight Newton tolerance,
plastic and bond-damage flags exactly equal.  The hierarchy is also checked for
what it promises: several levels, far fewer CG iterations than Jacobi on a
large plate, the single-level (dense coarsest) case, and bit-identical reruns.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import lpmg
from lpmg import lattice as lat
from lpmg import problems as P
from oracle import Oracle

MG64 = dict(precond=2, rtol=1e-24, atol=0.0, precision=0)
ORC = dict(precond=0, rtol=1e-24, atol=0.0)


def _relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _run(prob, backend, opts, tol=1e-11):
    s, gids = P.build(backend, prob)
    s.forward()
    stats = P.solve_problem(s, gids, P.Problem(**{**prob.__dict__, "tol": tol}), opts)
    return s, stats


@pytest.mark.parametrize("prob", [P.sc_cube(n=10, steps=3), P.hex_plate(nx=40, ny=36, steps=3),
                                  P.bcc_j2_bar(nx=3, ny=3, nz=6, steps=10, strain_step=5e-4)],
                         ids=["sc_cube", "hex_plate", "bcc_j2_bar"])
def test_multigrid_quasi_static_matches_oracle(prob):
    g, _ = _run(prob, lpmg.System, MG64)
    o, _ = _run(prob, Oracle, ORC)
    assert len(g.multigrid_levels()) >= 2
    x0 = o.get("XYZ_INITIAL")
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8
    assert _relerr(g.get("PIN"), o.get("PIN")) < 1e-8
    if prob.ptype == lat.J2PLASTICITY:
        ag, ao = g.get("STATE_VAR")[0::7], o.get("STATE_VAR")[0::7]
        np.testing.assert_array_equal(ag > 0, ao > 0)
        assert (ao > 0).any()


def test_multigrid_ct_with_bond_breaking_matches_oracle():
    prob = P.ct_plate(3000, 0.30, steps=12, opening=3e-5)
    g, _ = _run(prob, lpmg.System, MG64, tol=1e-10)
    o, _ = _run(prob, Oracle, ORC, tol=1e-10)
    x0 = o.get("XYZ_INITIAL")
    bg, bo = g.get("BDAMAGE") > 0.5, o.get("BDAMAGE") > 0.5
    assert bo.sum() > 0
    np.testing.assert_array_equal(bg, bo)
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8


def test_multigrid_fp32_inner_solve_matches_fp64_oracle():
    prob = P.hex_plate(nx=40, ny=36, steps=3)
    g, _ = _run(prob, lpmg.System, dict(precond=2, rtol=1e-4, atol=0.0, precision=1))
    o, _ = _run(prob, Oracle, ORC)
    x0 = o.get("XYZ_INITIAL")
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8


def _one_solve(prob, opts):
    s, gids = P.build(lpmg.System, prob)
    s.forward()
    P.apply_step(s, gids, prob.steps[0])
    s.assemble(precision=opts.get("precision", 0))
    s.forward()
    s.update_rr()
    x0 = s.get("XYZ")
    st = s.solve_linear(**opts)
    return s, int(st[0]), s.get("XYZ") - x0


def test_multigrid_cuts_cg_iterations():
    prob = P.hex_plate(nx=150, ny=150, steps=1)
    s1, it_j, du_j = _one_solve(prob, dict(precond=1, rtol=1e-20, atol=0.0))
    s2, it_m, du_m = _one_solve(prob, dict(precond=2, rtol=1e-20, atol=0.0))
    levels = s2.multigrid_levels()
    assert len(levels) >= 3 and levels[0] == 2 * 150 * 150 and levels[-1] <= 128
    assert it_m * 5 < it_j, (it_m, it_j)
    assert _relerr(du_m, du_j) < 1e-5


def test_single_level_dense_coarsest():
    prob = P.hex_plate(nx=8, ny=8, steps=2)  # 128 unknowns: the coarsest level is the system
    s, it, du = _one_solve(prob, dict(precond=2, rtol=1e-20, atol=0.0))
    assert s.multigrid_levels() == [128]
    assert it <= 2
    _, _, du_j = _one_solve(prob, dict(precond=1, rtol=1e-20, atol=0.0))
    assert _relerr(du, du_j) < 1e-10


def test_multigrid_reruns_are_bit_identical():
    prob = P.hex_plate(nx=60, ny=50, steps=3)
    a, _ = _run(prob, lpmg.System, MG64)
    b, _ = _run(prob, lpmg.System, MG64)
    np.testing.assert_array_equal(a.get("XYZ"), b.get("XYZ"))
