"""The mixed-precision workloads at their full sizes, checked through
properties that do not need the CPU oracle at that size: the FP32-inner path
against the FP64 GPU path (itself held to the oracle at smaller sizes by the
parity tests), exact plastic flags at every load step, and a request served
next to others giving the same bits as alone.

  BCC J2 bar: 30x30x60 cells (108,000 particles), E = 200 GPa, nu = 0.3,
              sigma_y = 250 MPa, H = 2 GPa, 45 tension steps of 2.5e-4
  polycrystal: BCC 80^3 cells, 200 Voronoi grains (about 1.0M particles),
              cubic C11/C12/C44 = 230/135/117 GPa, 10 tension steps of 1e-4,
              served together with CT requests

Both run the mixed path with Jacobi and with smoothed-aggregation multigrid
(the server's choice for BCC requests above its fusion limit).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import lpmg
from lpmg import problems as P
from lpmg import solver_opts
from lpmg.serve import Request, Server

FP64 = dict(precond=1, rtol=1e-12, atol=0.0, precision=0)
MIXED = dict(precond=1, rtol=1e-4, atol=0.0, precision=1)
# the server's choice for these BCC requests: smoothed-aggregation multigrid
MIXED_SA = dict(precond=4, rtol=1e-4, atol=0.0, precision=1)
NEWTON_TOL = 1e-11


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_bcc_j2_bar_mixed_matches_fp64_every_step():
    prob = P.bcc_j2_bar(30, 30, 60, steps=45, strain_step=2.5e-4)
    assert len(prob.xyz) == 108000
    s, gids = P.build(lpmg.System, prob)
    s.forward()
    runs = {}
    for name, o in (("fp64", FP64), ("mixed", MIXED), ("mixed_sa", MIXED_SA)):
        s.restart()
        us, flags = [], []
        for step in prob.steps:
            P.apply_step(s, gids, step)
            ok, _ = s.solve_step(prob.undamaged, prob.max_iter, NEWTON_TOL, **o)
            assert ok
            us.append(s.get("XYZ") - s.get("XYZ_INITIAL"))
            flags.append(s.get("STATE_VAR")[0::7] > 0)
        runs[name] = (us, flags)
    for name in ("mixed", "mixed_sa"):
        for k, (a, b) in enumerate(zip(runs[name][0], runs["fp64"][0])):
            assert _rel(a, b) <= 1e-8, (name, k)
        for fa, fb in zip(runs[name][1], runs["fp64"][1]):
            np.testing.assert_array_equal(fa, fb)
    assert runs["fp64"][1][-1].mean() > 0.9  # the bar is yielding throughout


def _req(p, o):
    sizes = {"BDAMAGE": len(p.extra["bonds"][0])} if "bonds" in p.extra else {}
    out = ("XYZ", "BDAMAGE") if sizes else ("XYZ",)
    return Request(p, solver_opts(**o), out_fields=out, sizes=sizes)


@pytest.fixture(scope="module")
def poly():
    p = P.bcc_polycrystal(80, 200, steps=10, strain_step=1e-4)
    p.tol = NEWTON_TOL
    assert len(p.xyz) > 1_000_000
    return p, {}


def test_polycrystal_mixed_matches_fp64(poly):
    p, alone = poly
    # with a fusion limit (the bench's 10k particles) the server picks
    # smoothed aggregation for this BCC request
    auto = dict(MIXED, precond=-1)
    for name, o in (("fp64", FP64), ("mixed", MIXED), ("mixed_auto", auto)):
        r = _req(p, o)
        with Server(max_active=1, batching=10000 if name == "mixed_auto" else True) as srv:
            status, st = srv.wait(srv.submit(r))
        assert status == 0, st.get("error")
        assert st["bytes"] <= st["bytes_est"]
        alone[name] = r.out["XYZ"].reshape(-1, 3) - p.xyz
        if name == "mixed_auto":
            assert st["cg"] < 2000, st["cg"]  # Jacobi takes 16,856 iterations
    assert _rel(alone["mixed"], alone["fp64"]) <= 1e-8
    assert _rel(alone["mixed_auto"], alone["fp64"]) <= 1e-8


def test_polycrystal_served_with_ct_requests(poly):
    p, alone = poly
    if "mixed" not in alone:
        r = _req(p, MIXED)
        with Server(max_active=1) as srv:
            assert srv.wait(srv.submit(r))[0] == 0
        alone["mixed"] = r.out["XYZ"].reshape(-1, 3) - p.xyz
    cts = [P.ct_plate(n, a, steps=30, opening=1e-5, tol=1e-9) for n, a in ((6000, 0.3), (20000, 0.45))]
    auto = dict(MIXED, precond=-1)
    ct_alone = []
    for c in cts:
        r = _req(c, auto)
        with Server(max_active=1, batching=50000) as srv:
            assert srv.wait(srv.submit(r))[0] == 0
        ct_alone.append(r)
    mix = [_req(p, MIXED)] + [_req(c, auto) for c in cts]
    with Server(max_active=4, batching=50000) as srv:
        res = [srv.wait(t) for t in [srv.submit(r) for r in mix]]
        info = srv.info()
    assert all(rc == 0 for rc, _ in res)
    assert info["peak_running"] >= 2
    np.testing.assert_array_equal(mix[0].out["XYZ"].reshape(-1, 3) - p.xyz, alone["mixed"])
    for a, b in zip(mix[1:], ct_alone):
        np.testing.assert_array_equal(a.out["XYZ"], b.out["XYZ"])
        np.testing.assert_array_equal(a.out["BDAMAGE"], b.out["BDAMAGE"])


# CT requests of the serving stream's workload (2D hex CT, a/W in [0.2, 0.5],
# critical bond stretch 1e-3, 30 load steps of pin opening 7.5e-6 W, Newton tol 1e-9)
CT_TOL = 1e-9
SERVED = dict(precond=-1, rtol=1e-5, atol=0.0, precision=1, predict=1.0)


def test_large_served_ct_matches_plain_fp64():
    # A 100k-particle CT through the server's mixed path (multigrid CG, FP32
    # inner solves, predictor) against the plain FP64 path on its own handle
    # (Jacobi-PCG at the reference's CG tolerance, the reference's step
    # start): displacements within the stream's bar, the same broken bonds.
    prob = P.ct_plate(100500, 0.29, steps=30, opening=7.5e-6, crit=1e-3, tol=CT_TOL)
    assert len(prob.xyz) >= 100000
    s, gids = P.build(lpmg.System, prob)
    s.forward()
    P.solve_problem(s, gids, prob, FP64)
    u64 = (s.get("XYZ") - s.get("XYZ_INITIAL")).reshape(-1, 3)
    bd64 = s.get("BDAMAGE") > 0.5
    s.close()
    r = _req(prob, SERVED)
    with Server(max_active=2, batching=50000) as srv:
        status, st = srv.wait(srv.submit(r))
    assert status == 0, st.get("error")
    assert st["bytes"] <= st["bytes_est"]
    u = r.out["XYZ"].reshape(-1, 3) - prob.xyz
    assert _rel(u, u64) <= 1e-8
    assert bd64.sum() > 0
    np.testing.assert_array_equal(r.out["BDAMAGE"] > 0.5, bd64)


def test_stream_sized_ct_matches_oracle():
    # A CT of the stream's smallest size class (>= 5k particles) served with
    # the mixed options against the CPU oracle's FP64 solve of the same
    # request: displacements within 1e-8, the same broken bonds.
    from oracle import Oracle, set_threads

    prob = P.ct_plate(5200, 0.35, steps=30, opening=7.5e-6, crit=1e-3, tol=CT_TOL)
    assert len(prob.xyz) >= 5000
    set_threads(16)
    o, gids = P.build(Oracle, prob)
    o.forward()
    P.solve_problem(o, gids, prob, dict(precond=1, rtol=1e-12, atol=0.0))
    u_ref = (o.get("XYZ") - o.get("XYZ_INITIAL")).reshape(-1, 3)
    bd_ref = o.get("BDAMAGE") > 0.5
    r = _req(prob, SERVED)
    with Server(max_active=1, batching=50000) as srv:
        status, st = srv.wait(srv.submit(r))
    assert status == 0, st.get("error")
    u = r.out["XYZ"].reshape(-1, 3) - prob.xyz
    assert _rel(u, u_ref) <= 1e-8
    assert bd_ref.sum() > 0
    np.testing.assert_array_equal(r.out["BDAMAGE"] > 0.5, bd_ref)
