"""The load-step predictor (solver option 6, lpmg_solve_step): a step starts
with the free DOFs moved by the last converged step's increment.  It changes
where Newton starts, not where it ends: results stay within 1e-8 of the
oracle's (which starts from the reference's point), damage and plastic flags
stay identical, and on the uniform elastic steps Newton needs fewer
iterations."""
import numpy as np
import pytest

import lpmg
from lpmg import lattice as lat
from lpmg import problems as P
from lpmg.serve import Request, Server
from oracle import Oracle

pytestmark = pytest.mark.gpu

GPU_OPTS = dict(precond=1, rtol=1e-24, atol=0.0)
ORC_OPTS = dict(precond=0, rtol=1e-24, atol=0.0)


def _relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _run(prob, backend, opts, tol=1e-11):
    s, gids = P.build(backend, prob)
    s.forward()
    stats = P.solve_problem(s, gids, P.Problem(**{**prob.__dict__, "tol": tol}), opts)
    return s, stats


@pytest.mark.parametrize("prob", [P.sc_cube(n=8, steps=6),
                                  P.bcc_j2_bar(nx=3, ny=3, nz=6, steps=10, strain_step=5e-4),
                                  P.ct_plate(2000, 0.30, steps=12, opening=3e-5)],
                         ids=["sc_cube", "bcc_j2_bar", "ct_damage"])
def test_predicted_steps_match_oracle(prob):
    # the CT's Newton residual floors near 1.6e-10 of its reaction (FP64
    # roundoff in absolute coordinates), so it converges to 1e-10
    tol = 1e-10 if prob.name.startswith("ct") else 1e-11
    g, st_g = _run(prob, lpmg.System, dict(GPU_OPTS, predict=1.0), tol)
    n, st_n = _run(prob, lpmg.System, GPU_OPTS, tol)
    o, _ = _run(prob, Oracle, ORC_OPTS, tol)
    x0 = o.get("XYZ_INITIAL")
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8
    assert _relerr(g.get("PIN"), o.get("PIN")) < 1e-8
    np.testing.assert_array_equal(g.get("BDAMAGE") > 0.5, o.get("BDAMAGE") > 0.5)
    if prob.ptype == lat.J2PLASTICITY:
        np.testing.assert_array_equal(g.get("STATE_VAR")[0::7] > 0, o.get("STATE_VAR")[0::7] > 0)
    if prob.name == "sc_cube":
        assert sum(s[2] for s in st_g) < sum(s[2] for s in st_n)  # fewer Newton iterations


def test_predictor_follows_halved_steps_and_serving():
    # a served request with the predictor equals the same solve through the
    # binding with the predictor, bit for bit, and the oracle within 1e-8
    prob = P.sc_cube(n=8, steps=4)
    prob.tol = 1e-11
    opts = dict(GPU_OPTS, predict=1.0)
    g, _ = _run(prob, lpmg.System, opts)
    req = Request(prob, lpmg.solver_opts(**opts), out_fields=("XYZ",))
    with Server(max_active=2, batching=True) as srv:
        rc, st = srv.wait(srv.submit(req))
    assert rc == 0
    np.testing.assert_array_equal(req.out["XYZ"], g.get("XYZ"))
    o, _ = _run(prob, Oracle, ORC_OPTS)
    x0 = o.get("XYZ_INITIAL")
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8
