"""Serving runtime on the GPU (lpmg_server_*): continuous batching of uneven
quasi-static solves with memory-gated admission.

Every request's arithmetic is the same alone and in a mix -- a fused launch
gives each system the workgroups and the reduction order its own launch would
-- so results are compared bit for bit: displacements, bond damage flags and
the Newton / CG / damage-pass counts.  One request is also checked against the
CPU oracle at the 1e-8 relative L2 bar with exact damage flags.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from lpmg import LpmgError, System, solver_opts
from lpmg import problems as P
from lpmg.serve import Request, Server

FP64 = dict(precond=1, rtol=1e-12, atol=0.0, precision=0)
MIXED = dict(precond=1, rtol=1e-4, atol=0.0, precision=1)
OUT = ("XYZ", "BDAMAGE", "DAMAGE")


def _opts(o):
    return solver_opts(**o)


def _problems():
    ct_a = P.ct_plate(2000, 0.30, steps=12, opening=3e-5)
    ct_b = P.ct_plate(5000, 0.45, steps=12, opening=3e-5)
    ct_c = P.ct_plate(3000, 0.25, steps=12, opening=3e-5)
    cube = P.sc_cube(n=8, steps=3, total=0.003)
    plate = P.hex_plate(40, 30, steps=3, total=0.003)
    return [(ct_a, FP64), (ct_b, FP64), (ct_c, MIXED), (cube, FP64), (plate, MIXED)]


def _direct(prob, o):
    """The per-handle path: lpmg.System driven by problems.solve_problem."""
    s, gids = P.build(System, prob)
    P.prepare(s)
    stats = P.solve_problem(s, gids, prob, o)
    out = dict(XYZ=s.get("XYZ"), BDAMAGE=s.get("BDAMAGE"), DAMAGE=s.get("DAMAGE"))
    tot = np.sum(np.array(stats), axis=0)
    s.close()
    return out, dict(cg=tot[0], passes=tot[1], newton=tot[2], steps=len(stats))


def _request(prob, o):
    nb = len(prob.extra["bonds"][0]) if "bonds" in prob.extra else None
    sizes = {"BDAMAGE": nb} if nb else {"BDAMAGE": _nbond(prob)}
    return Request(prob, _opts(o), out_fields=OUT, sizes=sizes)


_NB = {}


def _nbond(prob):
    key = id(prob)
    if key not in _NB:
        s, _ = P.build(System, prob)
        _NB[key] = s.nbond
        s.close()
    return _NB[key]


def _same(req, res, ref, ref_stats):
    status, st = res
    assert status == 0, st.get("error")
    for k in OUT:
        assert np.array_equal(req.out[k], ref[k]), k
    assert st["cg"] == ref_stats["cg"] and st["newton"] == ref_stats["newton"]
    assert st["passes"] == ref_stats["passes"] and st["steps"] == ref_stats["steps"]


@pytest.fixture(scope="module")
def alone():
    probs = _problems()
    return probs, [_direct(p, o) for p, o in probs]


def test_mix_matches_alone(alone):
    """Five uneven requests (2D FP64, 2D mixed, 3D) admitted together: each
    result is bit-identical to the same solve alone on its own handle."""
    probs, refs = alone
    reqs = [_request(p, o) for p, o in probs]
    with Server(max_active=len(reqs), batching=True) as srv:
        tickets = [srv.submit(r) for r in reqs]
        results = [srv.wait(t) for t in tickets]
        info = srv.info()
    for req, res, (ref, rs) in zip(reqs, results, refs):
        _same(req, res, ref, rs)
    assert info["peak_running"] >= 2          # they did run concurrently
    assert info["max_batch"] >= 2             # and shared fused CG launches
    assert info["finished"] == len(reqs)


def test_alone_in_server_and_unbatched(alone):
    """A request alone in a batching server, and the mix in a server without
    batching (each CG on its own stream), give the same bits too."""
    probs, refs = alone
    with Server(max_active=1, batching=True) as srv:
        r = _request(*probs[1])
        _same(r, srv.wait(srv.submit(r)), *refs[1])
    reqs = [_request(p, o) for p, o in probs[:3]]
    with Server(max_active=3, batching=False) as srv:
        res = [srv.wait(t) for t in [srv.submit(q) for q in reqs]]
    for q, rr, (ref, rs) in zip(reqs, res, refs[:3]):
        _same(q, rr, ref, rs)


def test_memory_gated_admission(alone):
    """A budget that holds one request defers the others (never more than one
    running) and they all finish; a request larger than the budget is refused."""
    probs, refs = alone
    prob, o = probs[0]
    reqs = [_request(prob, o) for _ in range(3)]
    est = reqs[0].bytes()
    with Server(max_active=3, mem_budget=int(1.5 * est), batching=True) as srv:
        tickets = [srv.submit(r) for r in reqs]
        results = [srv.wait(t) for t in tickets]
        info = srv.info()
        big = _request(*probs[1])
        assert big.bytes() > 1.5 * est
        with pytest.raises(LpmgError, match="-3"):
            srv.submit(big)
    assert info["peak_running"] == 1 and info["peak_reserved"] <= 1.5 * est
    for r, res in zip(reqs, results):
        _same(r, res, *refs[0])
        assert res[1]["bytes"] <= res[1]["bytes_est"]
    assert max(res[1]["queue_ms"] for res in results) > 0


def test_served_request_matches_oracle():
    """A damaged CT request through the server vs the CPU oracle: positions
    within 1e-8 relative L2, broken-bond flags exactly equal."""
    from oracle import Oracle
    prob = P.ct_plate(2000, 0.30, steps=12, opening=3e-5, tol=1e-10)
    o, gids = P.build(Oracle, prob)
    P.prepare(o)
    P.solve_problem(o, gids, prob, dict(precond=1, rtol=1e-12, atol=0.0))
    x_ref, bd_ref = o.get("XYZ"), o.get("BDAMAGE")
    x0 = prob.xyz
    req = _request(prob, FP64)
    with Server(max_active=1) as srv:
        status, st = srv.wait(srv.submit(req))
    assert status == 0, st.get("error")
    u, u_ref = req.out["XYZ"].reshape(-1, 3) - x0, x_ref.reshape(-1, 3) - x0
    assert np.linalg.norm(u - u_ref) <= 1e-8 * np.linalg.norm(u_ref)
    assert (bd_ref > 0.5).sum() > 0
    assert np.array_equal(req.out["BDAMAGE"] > 0.5, bd_ref > 0.5)


def test_two_level_reduction_in_fused_launch():
    """A system with more than 4096 SpMV workgroups (two-level partial sums)
    next to a small one: both equal their solo runs."""
    big = P.hex_plate(380, 380, steps=1, total=0.0005)
    small = P.hex_plate(30, 30, steps=2, total=0.001)
    refs = [_direct(big, FP64), _direct(small, FP64)]
    reqs = [_request(big, FP64), _request(small, FP64)]
    with Server(max_active=2) as srv:
        res = [srv.wait(t) for t in [srv.submit(q) for q in reqs]]
    for q, rr, (ref, rs) in zip(reqs, res, refs):
        _same(q, rr, ref, rs)


def test_server_picks_solver_by_size():
    """precond < 0: requests up to the fusion threshold run the fused Jacobi CG,
    larger 2D ones the smoothed-aggregation multigrid CG on their own stream --
    each bit-identical to the same solve on its own handle with that
    preconditioner."""
    small = P.ct_plate(2000, 0.30, steps=12, opening=3e-5)
    large = P.hex_plate(60, 60, steps=3, total=0.003)
    auto = dict(precond=-1, rtol=1e-4, atol=0.0, precision=1)
    ref_small = _direct(small, dict(auto, precond=1))
    ref_large = _direct(large, dict(auto, precond=4))
    reqs = [_request(small, auto), _request(large, auto)]
    with Server(max_active=2, batching=3000) as srv:
        res = [srv.wait(t) for t in [srv.submit(q) for q in reqs]]
        info = srv.info()
    _same(reqs[0], res[0], *ref_small)
    _same(reqs[1], res[1], *ref_large)
    assert info["chunks"] > 0
